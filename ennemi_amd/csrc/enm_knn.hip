// Dispatch of the kNN launches by joint dimension and list depth; the
// kernels themselves are in enm_knn.h, instantiated per dimension class
// by enm_knn_d*.hip.
#include "enm_knn.h"

int enm_knn_dim_class(int d) {
    if (d <= 6) return d;
    if (d <= 8) return 8;
    if (d <= 12) return 12;
    if (d <= 16) return 16;
    return -1;
}

int enm_knn_list_class(int k1) {
    if (k1 <= 4) return 4;
    if (k1 <= 8) return 8;
    if (k1 <= 16) return 16;
    if (k1 <= 32) return 32;
    if (k1 <= ENM_LIST_LDS_MAX) return ENM_LIST_HEAP_LDS;
    return ENM_LIST_HEAP_GLOBAL;
}

hipError_t enm_launch_knn(const Item* items, const int* qoff, int n_items, unsigned* counter,
                          const KnnDev* descs, const OrderDev* orders,
                          const unsigned long long* maxabs, const int* perm,
                          const double* col64, const float* col32, double* res64,
                          unsigned long long* stats, int dclass, int pclass, bool f32, bool brute,
                          int walk, int grid, int k1max, double* heap, int ticket, int smax, char* far,
                          unsigned* far_count, int64_t far_bytes, hipStream_t st) {
    if (n_items <= 0) return hipSuccess;
    // Tickets of up to ENM_KCHUNK items (or the tuned items per ticket), but
    // at least four per wave of the grid, so small launches still fill it.
    const int chunk = std::max(1, std::min(ticket > 0 ? ticket : ENM_KCHUNK, n_items / (4 * std::max(1, grid))));
    KnnArgs a{items, qoff, counter, chunk, descs, orders, maxabs, perm, brute ? 1 : 0,
              heap, (int64_t)ENM_WAVE * k1max, smax, far, far_count, far_bytes};
    if (pclass == ENM_LIST_HEAP_GLOBAL && heap == nullptr) return hipErrorInvalidValue;
    switch (dclass) {
        case 1: return enm_knn_launch_d1(a, pclass, f32, walk, grid, col64, col32, res64, stats, st);
        case 2: return enm_knn_launch_d2(a, pclass, f32, walk, grid, col64, col32, res64, stats, st);
        case 3: return enm_knn_launch_d3(a, pclass, f32, walk, grid, col64, col32, res64, stats, st);
        case 4: return enm_knn_launch_d4(a, pclass, f32, walk, grid, col64, col32, res64, stats, st);
        case 5: return enm_knn_launch_d5(a, pclass, f32, walk, grid, col64, col32, res64, stats, st);
        case 6: return enm_knn_launch_d6(a, pclass, f32, walk, grid, col64, col32, res64, stats, st);
        case 8: return enm_knn_launch_d8(a, pclass, f32, walk, grid, col64, col32, res64, stats, st);
        case 12: return enm_knn_launch_d12(a, pclass, f32, walk, grid, col64, col32, res64, stats, st);
        case 16: return enm_knn_launch_d16(a, pclass, f32, walk, grid, col64, col32, res64, stats, st);
        default: return hipErrorInvalidValue;
    }
}

int enm_knn_heap_grid(int grid, int k1max) {
    const int64_t per_block = (int64_t)ENM_WAVE * k1max * (int64_t)sizeof(double);
    const int64_t most = std::max<int64_t>(1, ENM_HEAP_GLOBAL_BYTES / std::max<int64_t>(1, per_block));
    return (int)std::max<int64_t>(1, std::min<int64_t>(grid, most));
}
