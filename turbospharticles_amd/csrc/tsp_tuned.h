// tsp_tuned.h -- build-time variant chosen by turbospharticles_amd/tune.py on MI355X
// (1M rigid cube drop: 3.662 ms per step; every candidate checked bit-identical). The halo cap
// 1064, lane sorting and u16 halo slots were then chosen on the window and the whole 500-step drop
// (the late steps of the drop have 5-12% of blocks with more than 1000 distinct FF neighbors).
#ifndef TSP_TUNE_BOX_CELLS
#define TSP_TUNE_BOX_CELLS 0
#endif
#ifndef TSP_TUNE_GATHER4
#define TSP_TUNE_GATHER4 0
#endif
#ifndef TSP_TUNE_GRAPHS
#define TSP_TUNE_GRAPHS 1
#endif
#ifndef TSP_TUNE_HALO_MAX
#define TSP_TUNE_HALO_MAX 1064
#endif
#ifndef TSP_TUNE_HALO_PAGES
#define TSP_TUNE_HALO_PAGES 0
#endif
#ifndef TSP_TUNE_JACOBI_REC24
#define TSP_TUNE_JACOBI_REC24 1
#endif
#ifndef TSP_TUNE_JACOBI_WAVES
#define TSP_TUNE_JACOBI_WAVES 6
#endif
#ifndef TSP_TUNE_KG_STORE
#define TSP_TUNE_KG_STORE 0
#endif
#ifndef TSP_TUNE_PA_WAVES
#define TSP_TUNE_PA_WAVES 6
#endif
#ifndef TSP_TUNE_PERSISTENT
#define TSP_TUNE_PERSISTENT 0
#endif
#ifndef TSP_TUNE_PREFETCH
#define TSP_TUNE_PREFETCH 1
#endif
#ifndef TSP_TUNE_QUERY_WAVES
#define TSP_TUNE_QUERY_WAVES 0
#endif
#ifndef TSP_TUNE_RADIX_BITS
#define TSP_TUNE_RADIX_BITS 8
#endif
#ifndef TSP_TUNE_RANGE_LOADS
#define TSP_TUNE_RANGE_LOADS 0
#endif
#ifndef TSP_TUNE_SEARCH_BATCH
#define TSP_TUNE_SEARCH_BATCH 4
#endif
#ifndef TSP_TUNE_SEARCH_ROW
#define TSP_TUNE_SEARCH_ROW 1
#endif
#ifndef TSP_TUNE_SLICE
#define TSP_TUNE_SLICE 64
#endif
#ifndef TSP_TUNE_STEP_GRAPH
#define TSP_TUNE_STEP_GRAPH 0
#endif
#ifndef TSP_TUNE_SWEEP_AHEAD
#define TSP_TUNE_SWEEP_AHEAD 2
#endif
#ifndef TSP_TUNE_SWEEP_THREADS
#define TSP_TUNE_SWEEP_THREADS 256
#endif
#ifndef TSP_TUNE_LANE_SORT
#define TSP_TUNE_LANE_SORT 1
#endif
#ifndef TSP_TUNE_SLOT16
#define TSP_TUNE_SLOT16 1
#endif
#ifndef TSP_TUNE_SEARCH_FLAT
#define TSP_TUNE_SEARCH_FLAT 0
#endif
