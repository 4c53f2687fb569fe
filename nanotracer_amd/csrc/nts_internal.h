// Internal definitions shared by the gfx950 translation units of libnts.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>
#include <string>
#include <vector>

#include "../../include/nts.h"
#include "nts_math.h"

// ---- parameter planes ------------------------------------------------------
// A scene stores every per-splat attribute as its own plane of `cap` floats
// (structure of arrays), so each kernel reads each attribute with fully
// coalesced 256-byte wave loads.  Gradients and both Adam moments use the
// same plane layout.  59 live floats per splat: no vec4 padding.
enum : int {
  P_POS = 0,     // 3 planes x, y, z
  P_ROT = 3,     // 4 planes w, x, y, z
  P_SCALE = 7,   // 3 planes log sigma
  P_LOGIT = 10,  // 1 plane
  P_DC = 11,     // 3 planes r, g, b
  P_REST = 14,   // 45 planes, planar R[15] G[15] B[15]
  P_COUNT = 59,
};

// Projected record (48 B, the ProjectedSplat layout):
// {mean.x, mean.y, depth, radius}{conic.a, conic.b, conic.c, visible}{r, g, b, opacity}
struct __align__(16) Projected {
  float4 a, b, c;
};
// Projected gradient (48 B, the ProjectedGrad layout):
// {dmean.x, dmean.y, dopacity, 0}{dconic.a, .b, .c, 0}{dr, dg, db, 0}
struct __align__(16) ProjectedGrad {
  float4 a, b, c;
};
// The batched step keeps its per-(view, splat) gradients in compact records
// of 9 words, the ProjectedGrad terms without the padding words:
// {dmean.x, dmean.y, dopacity, dconic.a, .b, .c, dr, dg, db}.
#define NTS_GRAD_WORDS 12u
#define NTS_GRAD9_WORDS 9u

struct DevCamera {  // the 96-byte camera, passed by value to kernels
  float view[16];
  float cam_pos[3];
  float near_;
  float focal;
  float width;
  float height;
  uint32_t n_splats;
};
static_assert(sizeof(DevCamera) == 96, "camera layout");
static_assert(sizeof(nts_camera) == 96, "camera layout");
static_assert(sizeof(nts_view_desc) == 112, "view layout");

struct ProfEntry {
  std::string name;
  uint64_t launches = 0;
  double total_ms = 0.0;
};

struct PendingEvent {
  int entry;
  hipEvent_t start, stop;
};

struct nts_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t hbm_budget = 0;
  uint64_t allocated = 0;
  bool prof = false;
  std::vector<ProfEntry> prof_entries;
  std::vector<PendingEvent> pending;
  std::vector<hipEvent_t> event_pool;
  // reusable scratch (grown on demand, never shrunk)
  void* scratch = nullptr;
  uint64_t scratch_bytes = 0;
  // pinned host word for small readbacks
  uint32_t* host_word = nullptr;
};

struct nts_buffer {
  void* ptr = nullptr;
  uint64_t bytes = 0;
};

struct nts_scene {
  uint32_t cap = 0;
  uint32_t n = 0;
  float* params = nullptr;   // P_COUNT planes
  float* grads = nullptr;    // P_COUNT planes
  float* m = nullptr;        // P_COUNT planes
  float* v = nullptr;        // P_COUNT planes
  float* grad_acc = nullptr; // 1 plane
  void* base = nullptr;      // single allocation
  uint64_t bytes = 0;
  uint64_t gen = 0;          // creation number, unique per process (keys captured graphs)
  bool trainable = true;     // holds gradients, moments and grad_acc
};
// Next scene creation number.
uint64_t nts_scene_next_gen();
// Calls that read or write gradients, moments or grad_acc need a scene
// created with NTS_SCENE_TRAINABLE.
#define NTS_NEED_TRAINABLE(s)                                                               \
  do {                                                                                      \
    if (!(s)->trainable)                                                                    \
      return nts_fail(NTS_E_STATE, "scene holds parameters only (created without "          \
                                   "NTS_SCENE_TRAINABLE): no gradients, moments or grad_acc"); \
  } while (0)

struct nts_bins {
  uint64_t total_pairs = 0;
  uint64_t capacity = 0;
  uint32_t num_tiles = 0;
  uint32_t tiles_x = 0, tiles_y = 0;
  uint32_t tile_size = 16;
  uint32_t key_bytes = 4;         // 4: (tile << 16 | depth16) u32, 8: u64
  void* keys = nullptr;           // sorted
  uint32_t* payloads = nullptr;   // sorted
  uint2* ranges = nullptr;        // per tile
  uint32_t* order = nullptr;      // blend order of the tiles, most pairs first (nullptr: grid order)
  void* base = nullptr;
  uint64_t bytes = 0;
};

// ---- error handling ---------------------------------------------------------
int nts_fail(int code, const char* fmt, ...);
#define NTS_HIP_TRY(expr)                                                            \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) return nts_fail(NTS_E_HIP, "%s: %s (%s:%d)", #expr,        \
                                          hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)
#define NTS_CHECK_ARG(cond)                                                          \
  do {                                                                               \
    if (!(cond)) return nts_fail(NTS_E_INVALID, "invalid argument: %s (%s:%d)", #cond, \
                                 __FILE__, __LINE__);                                \
  } while (0)

// Device allocation charged against the context's HBM budget.
int nts_dev_alloc(nts_ctx* ctx, void** p, uint64_t bytes);
void nts_dev_free(nts_ctx* ctx, void* p, uint64_t bytes);
// Host SoA -> device planes (transposed on the host), synchronous.
int nts_upload_planes(nts_ctx* ctx, const nts_scene* s, float* dst_planes, const nts_splats_host* h);
// Scene with parameter planes only (render-only residents).
int nts_scene_alloc_params(nts_ctx* ctx, uint32_t capacity, nts_scene** out);
// Context scratch of at least `bytes` (contents undefined).
int nts_scratch(nts_ctx* ctx, uint64_t bytes, void** out);

// ---- profiled launches ------------------------------------------------------
int nts_prof_begin(nts_ctx* ctx, const char* name, PendingEvent* pe);
void nts_prof_end(nts_ctx* ctx, PendingEvent* pe);
int nts_prof_resolve(nts_ctx* ctx);

#define NTS_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                        \
  do {                                                                                 \
    PendingEvent _pe;                                                                  \
    int _on = nts_prof_begin((ctx), (name), &_pe);                                     \
    hipLaunchKernelGGL(kernel, grid, block, shmem, (ctx)->stream, __VA_ARGS__);        \
    hipError_t _le = hipGetLastError();                                                \
    if (_on) nts_prof_end((ctx), &_pe);                                                \
    if (_le != hipSuccess)                                                             \
      return nts_fail(NTS_E_HIP, "launch %s: %s", (name), hipGetErrorString(_le));     \
  } while (0)

static inline uint32_t nts_div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---- internal entry points shared across translation units ------------------
// Fill a device range with a 32-bit word (graph-safe replacement for memset).
int nts_fill_u32(nts_ctx* ctx, void* p, uint32_t value, uint64_t n_words);
// Exclusive u32 scan of n values (in -> out may alias); optional u64 total
// written to a device word.
int nts_scan_u32_dev(nts_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* d_total);
// Stable segmented LSD radix sort of (key, u32 value) pairs over the low
// key_bits; segment s is [seg_begin[s], seg_begin[s+1]) (device array).  The
// result ends in keys/vals, or in the alt buffers when *result_in_alt.
int nts_sort_pairs_dev(nts_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                       const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                       bool* result_in_alt);
int nts_sort_pairs32_dev(nts_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                         const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                         bool* result_in_alt);
// Workspace variants: no allocation and no host synchronisation (capturable).
uint64_t nts_scan_ws_bytes(uint64_t n);
int nts_scan_u32_ws(nts_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* d_total, void* ws);
uint64_t nts_sort_ws_bytes(uint32_t num_segs, uint64_t max_pairs, uint32_t key_bytes);
int nts_sort_ws_launch32(nts_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                         const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                         void* ws, bool* result_in_alt, uint32_t** err);
int nts_sort_ws_launch64(nts_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                         const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t key_bits,
                         void* ws, bool* result_in_alt, uint32_t** err);
// The binning sort of (tile << 16 | depth16) keys: the order of the full
// sort over bits [0, 16 + tile_bits), the tile ranges (ranges holds
// total_tiles entries, zero-filled by the caller; d_tile_base the first tile
// of each segment, num_segs + 1 entries) and the blend order of the tiles
// (order, total_tiles entries: most pairs first).  Keys are below 2^(16 + tile_bits)
// or all ones; with all-ones keys, tile ids stay below 2^tile_bits - 1.
// The launch variants take the full sort's workspace and do not
// synchronise; the _dev variants wait and check the look-back error.
int nts_bin_sort_ws_launch32(nts_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                             const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t tile_bits,
                             const uint32_t* d_tile_base, uint32_t total_tiles, uint2* ranges, uint32_t* order,
                             void* ws, bool* result_in_alt, uint32_t** err);
int nts_bin_sort_ws_launch64(nts_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                             const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t tile_bits,
                             const uint32_t* d_tile_base, uint32_t total_tiles, uint2* ranges, uint32_t* order,
                             void* ws, bool* result_in_alt, uint32_t** err);
int nts_bin_sort32_dev(nts_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                       const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t tile_bits,
                       const uint32_t* d_tile_base, uint32_t total_tiles, uint2* ranges, uint32_t* order,
                       bool* result_in_alt);
int nts_bin_sort64_dev(nts_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                       const uint64_t* d_seg_begin, uint32_t num_segs, uint64_t max_pairs, uint32_t tile_bits,
                       const uint32_t* d_tile_base, uint32_t total_tiles, uint2* ranges, uint32_t* order,
                       bool* result_in_alt);
// Smallest b >= 1 with 2^b > tiles: tile ids then stay below 2^b - 1, as the
// binning sort needs where culled pairs carry the all-ones key.
inline uint32_t nts_tile_bits_strict(uint64_t tiles) {
  uint32_t b = 1;
  while (b < 48 && (1ull << b) <= tiles) ++b;
  return b;
}
