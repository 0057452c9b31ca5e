// Tile autotuner: picks the colour-pass layout of a scene or of a batch's
// element types by timing candidates on the device, and keeps the winners in
// a process-wide cache per (device, element type, size bucket, form).
//
// The layout knobs are those of GatherTile: the pass mode (gather, gather with
// cached gradients, element tasks); for the gathers the lanes per vertex
// (256 / lanes vertices per workgroup) and the quadrature points per lane;
// for the element tasks the incidences per block (the elements a workgroup
// stages in LDS) and the block count below which a colour takes the latency
// form.  The reference fixes its layout per kernel version
// (src/GPU/VBD/GPU_VBD_FEM.cu:619-650); it has no tuner.
//
// A candidate is timed with HIP events on the engine's stream:
//   scene: one colour sweep (one VBD iteration, or one explicit force pass);
//   batch: the colour-pass launches of one element type in one batched step.
// The state the timing runs move (positions, previous positions, velocities)
// is saved first and put back afterwards, so tuning leaves the scenes as they
// were.  Candidates are timed once; the best third is timed twice more and
// the fastest time wins.
#pragma once

#include <cstdint>
#include <vector>

#include "oe_batch.hpp"
#include "oe_scene.hpp"

namespace oe {

enum TuneForm : int { TUNE_SCENE = 0, TUNE_BATCH = 1 };

struct TuneRecord {
    int device = 0;
    int element = 0;
    int bucket = 0;   // floor(log2(vertices)): of the scene, or of the batch's scenes of this type
    int form = TUNE_SCENE;
    GatherTile tile;
    double ms = 0.0;      // time of the winner (see above)
    int candidates = 0;   // candidates timed (0: taken from the cache)
};

int size_bucket(int64_t vertices);

// Every layout the kernels accept for this element type (tasks and cached
// gradients only for VBD).
std::vector<GatherTile> tile_candidates(Element el, bool vbd, bool tasks_possible);

// Set a scene's colour-pass layout (building the tables it needs) and drop
// its captured graph; the previous layout stays on failure.
void apply_scene_tile(Scene& sc, const GatherTile& t);

// Tune a VBD or explicit scene (or take the cached winner unless `force`) and
// apply the winner.  The scene must not belong to a batch or a server.
TuneRecord autotune_scene(Scene& sc, bool force);

// Tune every element type of a batch (cache as above) and apply the winners.
std::vector<TuneRecord> autotune_batch(Batch& b, bool force);

// Snapshot of the process-wide cache; clear empties it.
std::vector<TuneRecord> tune_cache_records();
void tune_cache_clear();

}  // namespace oe
