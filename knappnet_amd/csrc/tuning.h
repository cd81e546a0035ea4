// Engine settings per graph class, written by knappnet_amd/autotune.py from
// timed runs on gfx950:sramecc+:xnack-, 256 CUs (the search, candidates and timings it used are
// kept with the profiles). Regenerate on a new device with
// `python -m knappnet_amd.autotune`.
#pragma once

namespace netbc {

struct TuneEntry {
    int kind;           // 0 breadth-first, 1 integer lengths (Dial), 2 many small graphs
                        // (netbc_betweenness_batch: only sweep_block, the workgroup size)
    int directed;       // 0 undirected, 1 directed, -1 either
    int log2n_min;      // node-count class: floor(log2 n) in [min, max]
    int log2n_max;
    double deg_min;     // stored arcs per node in [min, max)
    double deg_max;
    int slots;          // 64-source batches per super-batch
    double alpha;       // bottom-up when frontier out-arcs * alpha > open in-arcs
    int sweep_group;    // entries per dependency-sweep task at most
    int sweep_block;    // threads per workgroup of the dependency sweep
    int step_block;     // threads per workgroup of the bottom-up step and hub tasks
    int hub_in;         // in-degree above which a node is a hub
    int hub_chunk;      // in-arcs per hub task
    int bucket_width;   // integer lengths: distances per bucket
    const char* tuned_on;
};

constexpr TuneEntry kTuning[] = {
    {0, 1, 0, 40, 0.0, 1e30, 64, 2.0, 4, 256, 256, 64, 512, 1, "R-MAT scale 18 directed, first 32,768 sources"},
    {0, 0, 0, 40, 0.0, 1e30, 64, 4.0, 32, 256, 512, 64, 512, 1, "R-MAT scale 22 undirected, 4,096 sampled sources"},
    {1, -1, 0, 40, 0.0, 1e30, 32, 1.0, 32, 256, 256, 64, 512, 1, "BA n=200,000 integer weights 1-16, first 8,192 sources"},
    {2, 0, 0, 40, 0.0, 1e30, 0, 1.0, 32, 512, 256, 64, 512, 1, "1,024 Watts-Strogatz graphs, n 50-2,000, k 6, p 0.1"},
};

}  // namespace netbc
