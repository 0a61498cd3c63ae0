// Host-side state of libfasttau (contexts, volumes, solve jobs).  Internal.
#pragma once
#include "fast_tau.h"
#include "tau_internal.h"

#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace fth {

// Mixed form: the reliable-update step is launched after every kMixedUpdateEvery iterations
// and at the end of every polling chunk.  A volume waiting for its update skips the iterations in
// between, so its trajectory does not depend on this interval.
constexpr int kMixedUpdateEvery = 4;

// The stencil kernels address each vector of a volume with 32-bit byte offsets, so an FP64
// vector (8 bytes per voxel) must stay below 4 GiB.
constexpr size_t kMaxSolveVoxels = ((size_t)1 << 29) - 1;

// Red-black batches launch only the tiles of unfinished volumes once those are at most half
// of a launch of at least this many tiles.
constexpr int kRbCompactMin = 512;
// A one-volume red-black launch set of at most this many tiles has its U blocks reduce the
// Q pass's partials into alpha themselves instead of a finalize launch: the blocks then fit
// about one dispatch round, so the reduction at their start costs less than the launch it
// saves (MI355X: 464 tiles at 256^3 -0.6% per step; 1280 tiles at 512^3 +2% per iteration).
constexpr int kRbFoldMaxTiles = 768;

struct Shape {
    int tx, ty, tz;
};

// Tile shape of the red-black passes: ty rows of kRbTX slots per tile, tz planes per tile,
// and the fused Y+Q pass's prefetch distance in planes.
struct RbShape {
    int ty, tz, ahead;
};

struct PinBuf {
    void* ptr = nullptr;
    size_t size = 0;
};

}  // namespace fth

struct fast_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0;
    size_t held = 0;
    std::string err;
    int verbose = 0;
    std::multimap<size_t, void*> freelist;     // cached device blocks by rounded size
    std::unordered_map<void*, size_t> live;    // blocks in use
    fth::PinBuf pin[4];
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    unsigned long long* zero_count = nullptr;
    double* alpha_scratch = nullptr;           // this device's copy of the area kernels' static scratch
    std::map<std::string, fth::Shape> tuned;   // autotuned tile shapes by size class
    std::map<std::string, fth::RbShape> rb_tuned;   // autotuned red-black tile shapes by size class
    fth::RbShape rb_forced{0, 0, 0};           // red-black shape forced by fast_ctx_set_rb_shape
    fth::Shape forced{0, 0, 0};                // tile shape forced by fast_ctx_set_tile_shape
    int chunk = 16;                            // Krylov iterations per polling round
    bool use_graphs = true;                    // replay polling rounds from a captured HIP graph
    size_t group_bytes = (size_t)320 << 20;    // red-black batches: state bytes per sub-batch (0 = one launch)
    int update_every = fth::kMixedUpdateEvery; // mixed form: iterations between reliable-update steps
    // mixed BiCGSTAB: reset r0 = r when r0 and r are this close to orthogonal (0: only by
    // Eigen's rule), and drop the correction when |r|^2 grows this much since the last update
    // (infinite: only a non-finite residual does)
    double mixed_restart_cos = 0.0;
    double mixed_blowup2 = HUGE_VAL;
    std::unordered_map<fast_volume*, int> volumes;   // live volumes, released with the context
    std::mutex mu;
};

struct fast_volume {
    fast_ctx* ctx = nullptr;
    int nx = 0, ny = 0, nz = 0;
    uint8_t* d_mask = nullptr;
    uint8_t* d_code = nullptr;
    uint32_t* d_labels = nullptr;   // cached CCL labels
    uint32_t num_labels = 0;
    int labels_bg = -1;             // background value the cached labels were computed for
    uint64_t zeros = 0;
};

namespace fth {

struct Job {
    fast_volume* vol = nullptr;
    fast_tau_params prm{};
    fast_tau_result* res = nullptr;
    double* conc = nullptr;
    double porosity = 0;
    double update_delta = 0.1, outer_tol = 0;
    int tz = 16;
    Shape shape{64, 8, 16};
};

// CCL on the volume's mask (ccl_api.cpp): labels cached on the volume per background value
// (numLabels = K + 1), and the 7-row per-face label table in device memory (caller frees).
int ensure_labels(fast_volume* v, uint8_t background, uint32_t* num_labels);
int face_table(fast_volume* v, uint32_t numLabels, uint8_t** out);

void set_error(fast_ctx* ctx, const char* fmt, ...);
void* dalloc(fast_ctx* ctx, size_t bytes);
void dfree(fast_ctx* ctx, void* p);
void* pinned(fast_ctx* ctx, int slot, size_t bytes);
bool hip_ok(fast_ctx* ctx, hipError_t e, const char* what);
template <class T>
ft::Coef<T> make_coef(int nx, int ny, int nz, int dir, double d_pore, double d_solid);
std::string shape_key(int nx, int ny, int nz, int precision);
Shape choose_shape(int nx, int ny, int nz, int precision, const fast_ctx* ctx);
void load_tuning(fast_ctx* ctx);
int solve_jobs(fast_ctx* ctx, std::vector<Job>& jobs);
// Red-black tile shape of a volume: forced (fast_ctx_set_rb_shape), else the tuned entry of
// its size class (key "rb_<x>_<y>_<z>"), else 4 rows, the planes rb_tz gives and 2 planes of
// prefetch.  $FAST_TAU_RB_TZ overrides the planes per tile.
RbShape rb_shape(const fast_ctx* ctx, int nx, int ny, int nz);
// Planes per tile for a volume with hx slots per row: enough tiles to fill the chip.
int rb_tz(int hx, int ny, int nz, int ty);
std::string rb_key(int nx, int ny, int nz);
bool rb_ty_compiled(int ty);
// Form of the red-black Y and Q passes (RbLaunch::staged): FAST_TAU_RB_STAGED = 0 (separate,
// Q reads six Y values), 1 (separate, Q stages the Y planes in LDS), default 2 (fused).
int rb_mode();

}  // namespace fth
