#include "oe_scene.hpp"

#include <stdexcept>

namespace oe {

Scene::Scene(Engine* e, HostMesh m, const SceneConfig& c) : eng(e), mesh(std::move(m)), cfg(c) {
    if (!element_supported(mesh.element)) throw std::invalid_argument("unsupported element type");
    if (mesh.nb_vertices() <= 0 || mesh.topo.empty()) throw std::invalid_argument("empty mesh");
    const int S = elem_nb_vertices(mesh.element);
    if (mesh.topo.size() % S != 0) throw std::invalid_argument("topology size is not a multiple of the element size");
    for (int32_t v : mesh.topo)
        if (v < 0 || v >= mesh.nb_vertices()) throw std::invalid_argument("topology index out of range");
    if (!(cfg.young > 0.f) || !(cfg.poisson > -1.f && cfg.poisson < 0.5f) || !(cfg.density > 0.f))
        throw std::invalid_argument("invalid material parameters");
    {
        // masses need only the element volumes: sum over q of |det J| w
        FemConstants fc;
        fc.S = S;
        fc.E = mesh.nb_elements();
        fc.elem_volume = element_volumes(mesh);
        masses = compute_fem_mass(mesh, fc, cfg.density, cfg.distribution);
        for (float mv : masses)
            if (!(mv > 0.f)) throw std::invalid_argument("vertex with zero mass (not used by any element)");
    }
    ps = std::make_unique<GPU_ParticleSystem>(eng, mesh.xyz, masses);
    switch (cfg.solver) {
        case SolverVBD: {
            auto* integ = new GPU_VBD(ps.get(), cfg.iterations, cfg.rho);
            integrator.reset(integ);
            auto* d = new GPU_VBD_FEM(eng, mesh.element, mesh, cfg.material, cfg.young, cfg.poisson, cfg.damping,
                                      static_cast<VBD_Version>(cfg.vbd_version), cfg.target_color);
            d->set_shuffle_seed(cfg.shuffle_seed);
            integ->add_dynamics(d);
            fem = vbd_fem = d;
            break;
        }
        case SolverXPBD: {
            auto* integ = new GPU_PBD(cfg.global_damping);
            integrator.reset(integ);
            auto* d = new GPU_PBD_FEM(eng, mesh.element, mesh, cfg.young, cfg.poisson, cfg.material,
                                      static_cast<XPBD_FEM_VERSION>(cfg.xpbd_version));
            integ->add_dynamics(d);
            fem = pbd_fem = d;
            break;
        }
        case SolverExplicit: {
            auto* integ = new GPU_SemiExplicit();
            integrator.reset(integ);
            auto* d = new GPU_FEM_Explicit(eng, mesh.element, mesh, cfg.young, cfg.poisson, cfg.material, cfg.damping);
            integ->add_dynamics(d);
            fem = explicit_fem = d;
            break;
        }
        default: throw std::invalid_argument("unknown solver");
    }
    integrator->stream = eng->stream;
    for (int d = 0; d < 3; ++d) integrator->gravity[d] = cfg.gravity[d];
}

Scene::~Scene() { drop_graph(); }

std::vector<int> Scene::select_plane(const float o[3], const float n[3]) const {
    std::vector<int> ids;
    const int N = mesh.nb_vertices();
    for (int i = 0; i < N; ++i) {
        const float dx = mesh.xyz[3 * i] - o[0], dy = mesh.xyz[3 * i + 1] - o[1], dz = mesh.xyz[3 * i + 2] - o[2];
        if (dx * n[0] + dy * n[1] + dz * n[2] > 0.f) ids.push_back(i);
    }
    return ids;
}

void Scene::fix_vertices(const std::vector<int>& ids) {
    if (fix) throw std::logic_error("scene already has a fixation");
    drop_graph();
    fix = new GPU_Fix_Constraint(eng, mesh.xyz, ids);
    // constraints run in insertion order: fixation first, box limit after
    // (Cuda_Constraint_Rigid_Controller.cpp:16-20)
    integrator->_constraints.insert(integrator->_constraints.begin(), std::unique_ptr<GPU_Dynamic>(fix));
}

void Scene::box_limit(const float pmin[3], const float pmax[3]) {
    drop_graph();
    integrator->add_constraint(new GPU_Box_Limit(eng->stream, pmin, pmax));
}

void Scene::run_substeps(float dt, int substeps) {
    const float sub_dt = dt / static_cast<float>(substeps);
    for (int i = 0; i < substeps; ++i) integrator->step(ps.get(), sub_dt);
}

void Scene::set_graph(bool on) {
    if (!on) drop_graph();
    use_graph = on;
}

void Scene::drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
}

void Scene::step(float dt, int substeps) {
    if (!(dt > 0.f) || substeps <= 0) throw std::invalid_argument("step needs dt > 0 and substeps > 0");
    const bool graph_ok = use_graph && !(vbd_fem && (vbd_fem->shuffle || vbd_fem->time_gather));
    if (!graph_ok) {
        run_substeps(dt, substeps);
    } else {
        if (!graph_exec || graph_dt != dt || graph_sub != substeps) {
            drop_graph();
            OE_HIP(hipStreamBeginCapture(eng->stream, hipStreamCaptureModeThreadLocal));
            try {
                run_substeps(dt, substeps);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(eng->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            OE_HIP(hipStreamEndCapture(eng->stream, &graph));
            OE_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
            graph_dt = dt;
            graph_sub = substeps;
        }
        OE_HIP(hipGraphLaunch(graph_exec, eng->stream));
    }
    OE_LAUNCH_CHECK();
    steps_done++;
}

size_t Scene::device_bytes() const {
    size_t b = ps->device_bytes() + fem->topo.bytes() + fem->jxv.bytes() + fem->dN.bytes() + fem->wq.bytes();
    if (vbd_fem) {
        b += vbd_fem->d_slot_vid.bytes() + vbd_fem->d_slot_nbo.bytes() + vbd_fem->d_own.bytes() +
             vbd_fem->p_new.bytes() + vbd_fem->ju.bytes() + vbd_fem->lat_tables.blocks.bytes();
        for (const auto& kv : vbd_fem->thr_tables) b += kv.second.blocks.bytes();
    }
    return b;
}

}  // namespace oe
