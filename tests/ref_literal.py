"""Literal Python restatements of the reference's host-side colouring, used to
pin the C oracle and the engine on small meshes (test infrastructure).

They follow include/Tools/Graph.h statement by statement with the same data
structures: adjacency as sorted lists with repeats (std::multiset), DSAT's
queue as a sorted container of (sat, deg, vertex) popped from the top
(std::set<Node, maxSat>), Greedy_SLF's std::priority_queue with lazy removal.
"""
from __future__ import annotations

import heapq

from sortedcontainers import SortedList


def build_primal(topo, S):
    """Graph::build_primal (Graph.h:36-54)."""
    n = int(max(int(v) for e in topo for v in e)) + 1
    adj = [[] for _ in range(n)]
    for e in topo:
        for j in range(S):
            for k in range(S):
                if k == j:
                    continue
                adj[int(e[j])].append(int(e[k]))
    for a in adj:
        a.sort()
    return adj


def build_dual_vertex(topo, S):
    """Graph::build_dual_vertex (Graph.h:56-85) with add_edge (:147-153)."""
    n = len(topo)
    nverts = int(max(int(v) for e in topo for v in e)) + 1
    owners = [[] for _ in range(nverts)]
    for eid, e in enumerate(topo):
        for i in range(S):
            owners[int(e[i])].append(eid)
    adj = [[] for _ in range(n)]
    for i in range(nverts):
        o = owners[i]
        for j in range(len(o)):
            for k in range(j + 1, len(o)):
                adj[o[j]].append(o[k])
                adj[o[k]].append(o[j])
    for a in adj:
        a.sort()
    return adj


def dsat(adj):
    """GraphColoration::DSAT (Graph.h:321-371)."""
    n = len(adj)
    used = [False] * n
    colors = [-1] * n
    degree = [len(a) for a in adj]
    adj_cols = [set() for _ in range(n)]
    q = SortedList((0, degree[v], v) for v in range(n))
    while q:
        _, _, vid = q.pop(-1)                  # begin() of a set ordered by '>'
        for v in adj[vid]:
            if colors[v] != -1:
                used[colors[v]] = True
        i = 0
        while i < len(used) and used[i]:
            i += 1
        for v in adj[vid]:
            if colors[v] != -1:
                used[colors[v]] = False
        colors[vid] = i
        for v in adj[vid]:
            if colors[v] == -1:
                q.remove((len(adj_cols[v]), degree[v], v))
                adj_cols[v].add(i)
                degree[v] -= 1
                q.add((len(adj_cols[v]), degree[v], v))
    return colors, max(colors) + 1


def color_reduction(adj, colors, target):
    """GraphReduction::ColorReduction (Graph.h:941-969)."""
    colors = list(colors)
    to_color = sorted(v for v in range(len(adj)) if colors[v] >= target)
    for vid in to_color:
        occ = [0] * target
        for a in adj[vid]:
            if colors[a] >= target:
                continue
            occ[colors[a]] += 1
        colors[vid] = occ.index(min(occ))
    return colors


def greedy_slf(adj):
    """GraphColoration::Greedy_SLF (Graph.h:237-277) + find_color_greedy (:909-930)."""
    n = len(adj)
    heap = [(len(adj[i]), i) for i in range(n)]
    heapq.heapify(heap)
    ids = [0] * n
    removed = [False] * n
    degree = [len(a) for a in adj]
    for i in range(n - 1, -1, -1):
        while heap and removed[heap[0][1]]:
            heapq.heappop(heap)
        v = heap[0][1]
        heapq.heappop(heap)
        removed[v] = True
        ids[i] = v
        for nb in adj[v]:
            if not removed[nb]:
                degree[nb] -= 1
                heapq.heappush(heap, (degree[nb], nb))
    colors = [-1] * n
    nb_color = 0
    for i in ids:
        used = {colors[a] for a in adj[i] if colors[a] != -1}
        avail = [c for c in range(nb_color) if c not in used]
        if not avail:
            colors[i] = nb_color
            nb_color += 1
        else:
            colors[i] = avail[0]
    return colors, nb_color
