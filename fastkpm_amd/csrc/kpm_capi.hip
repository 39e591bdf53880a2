// kpm_capi.hip -- implementation of the C-ABI in include/fastkpm_hip.h.
//
// Owns device memory and the HIP stream of one engine handle, validates every
// shape on the host before a launch, picks the kernel geometry, and sequences
// the per-step launches of the Chebyshev recursions.  The algorithms follow
// Engine_CPU (src/engine_cpu.cpp) and Engine_cuSPARSE (src/engine_cusparse.cpp)
// of the reference; the kernels are in kpm_kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>


#include "fastkpm_hip.h"
#include "kpm_dispatch.h"
#include "kpm_internal.h"
// this unit defines the non-template helper kernels
#define FKD_DEFINE_HELPERS
#include "kpm_kernels.h"
#include "kpm_dots.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) return fail(FKPM_E_HIP, "%s failed: %s (%s:%d)", #x,       \
                                      hipGetErrorString(e_), __FILE__, __LINE__);    \
  } while (0)

#define CHK(x)              \
  do {                      \
    int rc_ = (x);          \
    if (rc_ != 0) return rc_; \
  } while (0)

template <class T>
struct TypeId {
  typedef T type;
};

// One recorded kernel launch: entry, grid, block and the argument bytes.
struct LaunchRec {
  const void* fn = nullptr;
  dim3 grid, block;
  std::vector<size_t> offs;
  std::vector<unsigned char> args;
};

// While a recorder is active on this thread for a stream, launch() on that
// stream appends to it instead of launching (LaunchRecorder below replays the
// run directly or as a cached hipGraph).
struct RecSink {
  hipStream_t stream = nullptr;
  std::vector<LaunchRec>* recs = nullptr;
};
thread_local RecSink g_rec;

// Launch a kernel through its host stub with exactly its parameter types.
template <typename... P>
int launch(void (*fp)(P...), dim3 grid, dim3 block, hipStream_t st,
           typename TypeId<P>::type... args) {
  void* argv[] = {(void*)&args...};
  if (g_rec.recs != nullptr && g_rec.stream == st) {
    static constexpr size_t sz[] = {sizeof(P)...};
    static constexpr size_t al[] = {alignof(P)...};
    LaunchRec r;
    r.fn = (const void*)fp;
    r.grid = grid;
    r.block = block;
    size_t off = 0;
    for (size_t k = 0; k < sizeof...(P); k++) {
      off = (off + al[k] - 1) / al[k] * al[k];
      r.offs.push_back(off);
      off += sz[k];
    }
    r.args.resize(off);
    for (size_t k = 0; k < sizeof...(P); k++) memcpy(r.args.data() + r.offs[k], argv[k], sz[k]);
    g_rec.recs->push_back(std::move(r));
    return 0;
  }
  hipError_t e = hipLaunchKernel((const void*)fp, grid, block, argv, 0, st);
  if (e != hipSuccess) return fail(FKPM_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
  return 0;
}

int replay_direct(const std::vector<LaunchRec>& recs, hipStream_t st) {
  std::vector<void*> argv;
  for (const LaunchRec& r : recs) {
    argv.resize(r.offs.size());
    for (size_t k = 0; k < r.offs.size(); k++)
      argv[k] = const_cast<unsigned char*>(r.args.data() + r.offs[k]);
    hipError_t e = hipLaunchKernel(r.fn, r.grid, r.block, argv.data(), 0, st);
    if (e != hipSuccess) return fail(FKPM_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
  }
  return 0;
}

// Identity of a recorded run: every launch's entry, grid and argument bytes.
uint64_t rec_hash(const std::vector<LaunchRec>& recs) {
  uint64_t x = 1469598103934665603ull;  // FNV-1a
  auto mix = [&](const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; i++) x = (x ^ b[i]) * 1099511628211ull;
  };
  for (const LaunchRec& r : recs) {
    mix(&r.fn, sizeof(r.fn));
    const unsigned g[6] = {r.grid.x, r.grid.y, r.grid.z, r.block.x, r.block.y, r.block.z};
    mix(g, sizeof(g));
    mix(r.args.data(), r.args.size());
  }
  return x;
}

bool rec_equal(const std::vector<LaunchRec>& a, const std::vector<LaunchRec>& b) {
  if (a.size() != b.size()) return false;
  for (size_t k = 0; k < a.size(); k++) {
    const LaunchRec &x = a[k], &y = b[k];
    if (x.fn != y.fn || x.grid.x != y.grid.x || x.grid.y != y.grid.y || x.grid.z != y.grid.z ||
        x.block.x != y.block.x || x.block.y != y.block.y || x.block.z != y.block.z ||
        x.args != y.args)
      return false;
  }
  return true;
}

// Instantiated graphs of recorded runs, most recently used first.
struct GraphCache {
  struct Entry {
    uint64_t key;
    std::vector<LaunchRec> recs;
    hipGraphExec_t exec;
  };
  std::vector<Entry> entries;
  static constexpr size_t kMax = 16;
  void clear() {
    for (Entry& e : entries) (void)hipGraphExecDestroy(e.exec);
    entries.clear();
  }
  ~GraphCache() { clear(); }
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n) {
    if (n <= bytes && p != nullptr) return 0;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      bytes = 0;
      if (e != hipSuccess) return fail(FKPM_E_HIP, "hipFree: %s", hipGetErrorString(e));
    }
    if (n == 0) return 0;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(FKPM_E_HIP, "hipMalloc(%zu bytes): %s", n, hipGetErrorString(e));
    }
    bytes = n;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

size_t host_elem_bytes(int dtype) {
  switch (dtype) {
    case FKPM_F32: return 4;
    case FKPM_F64: return 8;
    case FKPM_C32: return 8;
    case FKPM_C64: return 16;
  }
  return 0;
}
bool dtype_complex(int dtype) { return dtype == FKPM_C32 || dtype == FKPM_C64; }
bool dtype_single(int dtype) { return dtype == FKPM_F32 || dtype == FKPM_C32; }

int grid_for(size_t n, int per_block = 256) {
  size_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (int)g;
}

// Validate a BSR/COO block pattern on the host.
int check_pattern(const char* what, int nb, int nnzb, const int* row_ptr, const int* row_idx,
                  const int* col_idx) {
  if (nb <= 0 || nnzb < 0) return fail(FKPM_E_ARG, "%s: bad sizes nb=%d nnzb=%d", what, nb, nnzb);
  if (!row_ptr || (nnzb > 0 && !col_idx))
    return fail(FKPM_E_ARG, "%s: null pattern array", what);
  if (row_ptr[0] != 0 || row_ptr[nb] != nnzb)
    return fail(FKPM_E_ARG, "%s: row_ptr[0]=%d row_ptr[nb]=%d, expected 0 and %d", what,
                row_ptr[0], row_ptr[nb], nnzb);
  for (int i = 0; i < nb; i++) {
    if (row_ptr[i + 1] < row_ptr[i]) return fail(FKPM_E_ARG, "%s: row_ptr not monotone at %d", what, i);
    for (int p = row_ptr[i]; p < row_ptr[i + 1]; p++) {
      if (col_idx[p] < 0 || col_idx[p] >= nb)
        return fail(FKPM_E_ARG, "%s: col_idx[%d]=%d out of range [0,%d)", what, p, col_idx[p], nb);
      if (row_idx && row_idx[p] != i)
        return fail(FKPM_E_ARG, "%s: row_idx[%d]=%d inconsistent with row_ptr (row %d)", what, p,
                    row_idx[p], i);
    }
  }
  return 0;
}

// Kernel kinds with a geometry of their own.
// K_MOMT / K_ADS: the forward step writing the table of forward vectors and
// the backward step reading it (tuned apart: a step that writes a fresh
// vector per launch meets the caches differently from one cycling a ring).
enum KKind { K_MOM = 0, K_XI = 1, K_AD = 2, K_OUT = 3, K_MOMT = 4, K_ADS = 5, NKIND = 6 };

// Geometry of one kernel kind: lanes per block row, packets per lane,
// workgroup cap and nontemporal bits; nwg/chunk derived from them.
struct KCfg {
  int L = 0, KQ = 0, maxwg = 0, nt = 0;
  int mfma = 0;  // matrix-core kernel (one wave per block row: L = 64, KQ = 1):
                 // 1 general, 2 row-resident (rows of <= 8 blocks, vectors < 4 GiB)
  int sweep = 0;  // rows: 0 per-XCD regions, s >= 1 one front in segments of s passes
                  // per XCD (chunk = -s, workgroups a multiple of 8 s)
  int dm = 8;     // row-resident matrix-core kernels: block slots per row (7 when no
                  // row has more than 7 blocks, else 8); set by set_kind
  int nwg = 0, chunk = 0;
};

// Most blocks per row of the row-resident matrix-core kernels (fkd::MF_DM).
constexpr int kRowResidentMax = 8;

template <class S>
fkd::RSource<S> no_rsource() {
  fkd::RSource<S> r;
  r.rd = nullptr;
  r.rcol = nullptr;
  r.rval = nullptr;
  r.cr = 0.0;
  return r;
}

}  // namespace

struct fkpm_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  int dtype = FKPM_C64;
  int prec = FKPM_PREC_FP64;
  bool cx = true;
  bool single = false;  // storage is FP32

  // --- H (scaled) ---
  int nb = 0, b = 0, nnzb = 0, max_deg = 0;
  bool h_set = false;
  DevBuf rp, ci, hv, rowblk, hraw;
  double avg = 0.0, mag = 1.0;
  std::vector<int> h_rp, h_ci;

  // --- R and recursion state ---
  int n = 0, s = 0;
  bool r_set = false, r_dirty = true;
  DevBuf rraw;             // R as given (host layout and type)
  bool r_compact = false;  // at most one nonzero per row (probing colours)
  DevBuf rcol, rval, rflag;
  int qp = 0, sp = 0, wmax = 0, geo_b = 0;
  KCfg kc[NKIND];
  fkd::KSet<double> kd[NKIND];
  fkd::KSet<float> kf[NKIND];
  int force_L = 0, force_KQ = 0, max_wg = 0;
  bool tuned = false;
  bool autotune_on_use = false;
  bool mfma_on = true;  // matrix-core kernels where the block shape has them
  bool rr_force = false;  // FKPM_OPT_MFMA 2: row-resident ones by default wherever they exist
  int stoch_method = FKPM_STOCH_AUTO;
  int stoch_used = -1;  // method the last stoch_matrix ran (FKPM_STOCH_*)
  DevBuf R, A[3], Bv[3], xi, partial, psum, tune_dacc;
  int alpha_M = 0;  // M of the moments whose final state is held, 0 if none
  DevBuf atab;            // alpha_1 .. alpha_{M/2} of the last moments(M), when it fits
  bool a_stored = false;  // the forward vectors live in atab (else the ring A[3])
  int tab_K = 0;          // atab holds alpha_1 .. alpha_tab_K of the current H and R
  DevBuf tabc;            // stoch_matrix weights of the table vectors
  std::vector<double> tabc_host;
  int store_alpha = 1;    // FKPM_OPT_STORE_ALPHA: 0 never, 1 auto, 2 whenever the table fits
  int sweep_force = -1;  // FKPM_OPT_SWEEP: row sweep of every kind (-1: per configuration)
  int slice_force = -1;  // FKPM_OPT_SLICE: column-pass loop order (-1: per configuration)
  int order_force = -1;  // FKPM_OPT_ROW_ORDER: lattice row order (-1: per configuration)
  bool has_order = false;  // rdesc holds a row order for the current pattern
  DevBuf rdesc;            // {row, first block, end block} per sweep position
  long long table_cap_mb = -1;  // FKPM_OPT_ALPHA_TABLE_MB (-1: the default budget)

  // --- D ---
  int nbD = 0, nnzbD = 0;
  bool d_set = false, d_sym = false, d_fused = false, d_valid = false;
  DevBuf drp, dci, tpos, dacc, dout;
  std::vector<int> hd_rp, hd_ci;

  // --- two-current moments (moments2) ---
  DevBuf j_rp[2], j_ci[2], j_rb[2], j_raw[2], j_val[2], m2_a, m2_b, m2_ring, m2_c, m2_part;

  // --- hipGraph replay of the step loops (short steps) ---
  int graph_mode = -1;  // FKPM_OPT_GRAPH: -1 automatic, 0 never, 1 always
  GraphCache gcache;
  long long n_graph_launches = 0, n_graph_captures = 0;

  // --- timing ---
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double ms_mom = 0, ms_stoch = 0, ms_ad = 0;
  int n_mom = 0, n_stoch = 0, n_ad = 0;

  ~fkpm_handle() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    DevBuf* bufs[] = {&rp,    &ci,    &hv,      &rowblk, &hraw, &rraw,  &rcol,     &rval,
                      &rflag, &R,     &A[0],    &A[1],   &A[2], &Bv[0], &Bv[1],    &Bv[2],
                      &xi,    &partial, &psum,  &tune_dacc, &drp, &dci,  &tpos,     &dacc,
                      &dout,  &atab,  &tabc};
    for (DevBuf* d : bufs) d->release();
    DevBuf* m2bufs[] = {&j_rp[0], &j_ci[0], &j_rb[0], &j_raw[0], &j_val[0], &j_rp[1], &j_ci[1],
                        &j_rb[1], &j_raw[1], &j_val[1], &m2_a, &m2_b, &m2_ring, &m2_c, &m2_part};
    for (DevBuf* d : m2bufs) d->release();
    gcache.clear();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  size_t vbytes() const { return (size_t)n * sp * (cx ? 2 : 1) * (single ? 4 : 8); }
  size_t accbytes() const { return (size_t)n * sp * (cx ? 2 : 1) * 8; }
  int ne() const { return (single ? 4 : 2) / (cx ? 2 : 1); }  // elements per 16-byte packet
  // forward-phase alpha_k: R, then the table of every alpha_k (a_stored) or
  // the ring of three
  void* vec_a(int k) {
    if (k == 0) return R.p;
    if (a_stored) return static_cast<char*>(atab.p) + (size_t)(k - 1) * vbytes();
    return A[k % 3].p;
  }
  template <class S>
  const fkd::KSet<S>& ks(int kind) const;
};

template <>
const fkd::KSet<double>& fkpm_handle::ks<double>(int kind) const {
  return kd[kind];
}
template <>
const fkd::KSet<float>& fkpm_handle::ks<float>(int kind) const {
  return kf[kind];
}

namespace {

// Vectors of at most this many scalars count as short steps: their launches,
// not HBM, set the time of a step loop, so the loop is replayed as a hipGraph.
constexpr size_t kGraphMaxScalars = (size_t)1 << 22;

// Records the kernel launches of one step loop on the handle's stream; run()
// then launches them directly or, for short steps (FKPM_OPT_GRAPH), replays
// the cached hipGraph of exactly the same launches, capturing it the first
// time.  A graph replays launches with identical entries, grids and argument
// bytes, so it computes what the direct launches compute.
class LoopRecorder {
 public:
  explicit LoopRecorder(fkpm_handle* h) : h_(h) {
    g_rec.stream = h->stream;
    g_rec.recs = &recs_;
  }
  ~LoopRecorder() { stop(); }
  int run() {
    stop();
    const bool graph = h_->graph_mode == 1 ||
                       (h_->graph_mode == -1 && recs_.size() >= 8 &&
                        (size_t)h_->n * h_->sp * (h_->cx ? 2 : 1) <= kGraphMaxScalars);
    if (!graph) return replay_direct(recs_, h_->stream);
    GraphCache& gc = h_->gcache;
    const uint64_t key = rec_hash(recs_);
    for (size_t k = 0; k < gc.entries.size(); k++) {
      if (gc.entries[k].key == key && rec_equal(gc.entries[k].recs, recs_)) {
        std::rotate(gc.entries.begin(), gc.entries.begin() + k, gc.entries.begin() + k + 1);
        HIPCHK(hipGraphLaunch(gc.entries[0].exec, h_->stream));
        h_->n_graph_launches++;
        return 0;
      }
    }
    HIPCHK(hipStreamBeginCapture(h_->stream, hipStreamCaptureModeThreadLocal));
    const int rc = replay_direct(recs_, h_->stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h_->stream, &g);
    if (rc != 0 || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc != 0 ? rc : fail(FKPM_E_HIP, "graph capture: %s", hipGetErrorString(e));
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(FKPM_E_HIP, "graph instantiate: %s", hipGetErrorString(ei));
    if (gc.entries.size() >= GraphCache::kMax) {
      HIPCHK(hipStreamSynchronize(h_->stream));  // no replay of the evicted graph in flight
      (void)hipGraphExecDestroy(gc.entries.back().exec);
      gc.entries.pop_back();
    }
    gc.entries.insert(gc.entries.begin(), GraphCache::Entry{key, std::move(recs_), ex});
    h_->n_graph_captures++;
    HIPCHK(hipGraphLaunch(ex, h_->stream));
    h_->n_graph_launches++;
    return 0;
  }

 private:
  void stop() {
    if (g_rec.recs == &recs_) {
      g_rec.recs = nullptr;
      g_rec.stream = nullptr;
    }
  }
  fkpm_handle* h_;
  std::vector<LaunchRec> recs_;
};

bool lookup_kset(fkpm_handle* h, int kind, int B, int L, int KQ) {
  if (h->single)
    return h->cx ? fkd::find_kset_fc(B, L, KQ, &h->kf[kind]) : fkd::find_kset_fr(B, L, KQ, &h->kf[kind]);
  return h->cx ? fkd::find_kset_dc(B, L, KQ, &h->kd[kind]) : fkd::find_kset_dr(B, L, KQ, &h->kd[kind]);
}

bool compiled(int b, int L, int KQ) {
  fkd::GeomChoice gs[64];
  const int ng = fkd::geometries_for_block(b, gs, 64);
  for (int k = 0; k < ng; k++)
    if (gs[k].L == L && gs[k].KQ == KQ) return true;
  return false;
}

void derive_grid(int nb, KCfg& c) {
  const int gwg = (fkd::WG / fkd::WAVE) * (fkd::WAVE / c.L);
  const int cap = c.maxwg > 0 ? c.maxwg : 4096;
  int nwg = (nb + gwg - 1) / gwg;
  if (nwg > cap) nwg = cap;
  if (nwg < 1) nwg = 1;
  int chunk = (nb + nwg - 1) / nwg;
  chunk = ((chunk + gwg - 1) / gwg) * gwg;
  c.nwg = (nb + chunk - 1) / chunk;
  c.chunk = chunk;
  if (c.sweep > 0) {
    // one front: a multiple of 8 s workgroups (the extra ones find no rows)
    const int q = 8 * c.sweep;
    c.nwg = (c.nwg + q - 1) / q * q;
    c.chunk = -c.sweep;
  }
}

// Matrix-core kernels exist for complex 8x8 blocks (forward/Clenshaw steps,
// also into the table, and the autodiff step that recomputes alpha).
bool mfma_avail(const fkpm_handle* h, int kind) {
  return h->mfma_on && h->cx && h->b == 8 && (kind == K_MOM || kind == K_MOMT || kind == K_AD);
}

// The row-resident variants: rows of at most kRowResidentMax blocks (MF_DM of
// kpm_mfma.h), byte offsets of a vector within 32 bits.
bool mfma_rr_avail(const fkpm_handle* h, int kind) {
  return mfma_avail(h, kind) && h->max_deg <= kRowResidentMax && h->vbytes() < (size_t(1) << 32);
}

int set_kind(fkpm_handle* h, int kind, KCfg c) {
  if (c.mfma) {
    if (!mfma_avail(h, kind) || (c.mfma == 2 && !mfma_rr_avail(h, kind)))
      return fail(FKPM_E_UNSUP, "no matrix-core kernel for block %d (kind %d)", h->b, kind);
    c.L = 64;  // one wave per block row
    c.KQ = 1;
    c.nt = 0;
    c.dm = h->max_deg <= 7 ? 7 : 8;
  }
  if (!lookup_kset(h, kind, h->b, c.L, c.KQ))
    return fail(FKPM_E_UNSUP, "no kernel for block %d, lanes %d, packets %d", h->b, c.L, c.KQ);
  if (h->sweep_force >= 0) c.sweep = h->sweep_force;
  if (h->slice_force >= 0 && !c.mfma)
    c.nt = (c.nt & ~fkd::NT_SLICE) | (h->slice_force ? fkd::NT_SLICE : 0);
  if (h->order_force >= 0 && !c.mfma && kind != K_OUT)
    c.nt = (c.nt & ~fkd::NT_ORDER) | (h->order_force ? fkd::NT_ORDER : 0);
  // the order exists for lattice patterns only; its sweep is one range per XCD
  if (!h->has_order || c.mfma || kind == K_OUT) c.nt &= ~fkd::NT_ORDER;
  if (c.nt & fkd::NT_ORDER) c.sweep = 0;
  derive_grid(h->nb, c);
  h->kc[kind] = c;
  return 0;
}

// The row order in the chosen configurations after the pattern or the
// FKPM_OPT_ROW_ORDER switch changed: off where the pattern has none, forced
// on or off by the switch, else as chosen (its sweep is one range per XCD).
void apply_row_order(fkpm_handle* h) {
  for (int k = 0; k < NKIND; k++) {
    KCfg& c = h->kc[k];
    if (c.L <= 0) continue;
    const bool can = h->has_order && !c.mfma && k != K_OUT;
    if (!can || h->order_force == 0) c.nt &= ~fkd::NT_ORDER;
    else if (h->order_force == 1) c.nt |= fkd::NT_ORDER;
    if (c.nt & fkd::NT_ORDER) c.sweep = 0;
    derive_grid(h->nb, c);
  }
}

// The row array a step kernel of configuration c reads: the row offsets, or
// the row order's descriptors (NT_ORDER).
const int* rp_for(const fkpm_handle* h, const KCfg& c) {
  return (c.nt & fkd::NT_ORDER) ? h->rdesc.as<const int>() : h->rp.as<const int>();
}

// Kernel entry of a kind under its configuration.
template <class S>
auto mom_fn(const fkpm_handle* h, const fkd::KSet<S>& ks, const KCfg& c) {
  if (c.mfma == 2 && ks.mfma_mom_rr2 != nullptr && h->qp % 16 == 0)
    return c.dm == 7 ? ks.mfma_mom_rr2_7 : ks.mfma_mom_rr2;
  return c.mfma == 2 ? (c.dm == 7 ? ks.mfma_mom_rr7 : ks.mfma_mom_rr) : c.mfma ? ks.mfma_mom : ks.mom;
}
// The row-resident autodiff step in FP32 takes column tiles in pairs when the
// padded width allows (16 packets, two 16-column tiles, divide it): C4 FP32
// 1747 -> 1561 us per step.
template <class S>
auto adf_fn(const fkpm_handle* h, const fkd::KSet<S>& ks, const KCfg& c) {
  if (c.mfma == 2 && ks.mfma_ad_rr2 != nullptr && h->qp % 16 == 0)
    return c.dm == 7 ? ks.mfma_ad_rr2_7 : ks.mfma_ad_rr2;
  return c.mfma == 2 ? (c.dm == 7 ? ks.mfma_ad_rr7 : ks.mfma_ad_rr)
                      : c.mfma ? ks.mfma_ad_fused : ks.ad_fused;
}
template <class S>
auto adp_fn(const fkd::KSet<S>& ks, const KCfg& c) {
  return c.mfma ? ks.mfma_ad_plain : ks.ad_plain;
}

// Default geometry of a kernel kind when nothing is tuned: one pass of width
// wmax; the single-gather kinds take the widest lane groups (wave-uniform
// rows), the kinds with block reductions groups of 16 lanes (one DPP row).
KCfg default_cfg(fkpm_handle* h, int kind) {
  static const int pref_gather[] = {64, 32, 16, 8, 4};
  static const int pref_reduce[] = {16, 8, 32, 4, 64};
  const int* pref = (kind == K_MOM || kind == K_XI || kind == K_MOMT) ? pref_gather : pref_reduce;
  KCfg c;
  for (int t = 0; t < 5 && c.L == 0; t++) {
    const int L = pref[t];
    if (h->wmax % L == 0 && compiled(h->b, L, h->wmax / L)) {
      c.L = L;
      c.KQ = h->wmax / L;
    }
  }
  if (c.L == 0) {  // any compiled geometry of width <= wmax, widest first
    fkd::GeomChoice gs[64];
    const int ng = fkd::geometries_for_block(h->b, gs, 64);
    for (int k = 0; k < ng; k++) {
      const int w = gs[k].L * gs[k].KQ;
      if (w <= h->wmax && (c.L == 0 || w > c.L * c.KQ)) {
        c.L = gs[k].L;
        c.KQ = gs[k].KQ;
      }
    }
  }
  c.maxwg = 4096;
  c.nt = 0;
  // one front in segments of two passes: what the tuner picks on the 2-D and
  // 3-D lattices measured (C2 FP64: forward step 91 -> 81 us, autodiff 188
  // -> 166 us against per-XCD regions)
  c.sweep = 2;
  if (mfma_avail(h, kind)) {
    c.mfma = 1;
    // row-resident: the forward step always, the autodiff step in FP64 and in
    // FP32 where its column tiles pair up (adf_fn)
    if (mfma_rr_avail(h, kind) && (kind == K_MOM || kind == K_MOMT || !h->single ||
                                   h->qp % 16 == 0 || h->rr_force))
      c.mfma = 2;
  }
  return c;
}

// Whether moments(M) keeps the forward vectors (when they fit): always with
// FKPM_OPT_STORE_ALPHA 2; with 1 (automatic) when the three-vector ring
// exceeds the 32 MB of L2: below that the forward steps lose more on the
// table's write-back than the backward steps gain.  (Measured at C2, FP32:
// forward step 35 -> 38 us on a table, autodiff 107 -> 91 us; FP64: 81 ->
// 81 us, 191 -> 167 us.)
bool table_wanted(const fkpm_handle* h) {
  return h->store_alpha == 2 || (h->store_alpha == 1 && 3 * h->vbytes() > ((size_t)32 << 20));
}

// Process-wide cache of tuned geometries, keyed by the problem shape.
std::mutex g_tune_mu;
std::map<std::string, std::array<KCfg, NKIND>> g_tune_cache;

std::string tune_key(const fkpm_handle* h) {
  char k[128];
  snprintf(k, sizeof(k), "%d/%d/%d/%d/%d/%d/%d/%d/%d/%d", h->dtype, h->prec, h->b, h->s, h->nb,
           h->nnzb, h->mfma_on ? (h->rr_force ? 2 : 1) : 0,
           h->max_deg <= 7 ? 2 : h->max_deg <= kRowResidentMax ? 1 : 0, table_wanted(h) ? 1 : 0,
           h->has_order ? 1 : 0);
  return k;
}

// Packed width: qp packets per row, a multiple of wmax, the smallest compiled
// pass width that covers the row (or the widest, with several passes).
int choose_geometry(fkpm_handle* h) {
  const int ne = h->ne();
  const int qp0 = (h->s + ne - 1) / ne;
  fkd::GeomChoice gs[64];
  const int ng = fkd::geometries_for_block(h->b, gs, 64);
  if (ng == 0) return fail(FKPM_E_UNSUP, "block size %d is not supported", h->b);
  int wmax = 0, widest = 0;
  for (int k = 0; k < ng; k++) {
    const int w = gs[k].L * gs[k].KQ;
    widest = std::max(widest, w);
    if (w >= qp0 && (wmax == 0 || w < wmax)) wmax = w;
  }
  if (wmax == 0) wmax = widest;
  // matrix-core tiles are 16 complex columns wide
  if (h->mfma_on && h->cx && h->b == 8) wmax = std::max(wmax, 16 / ne);
  if (h->force_L > 0) {
    const int fkq = h->force_KQ > 0 ? h->force_KQ : 1;
    if (!compiled(h->b, h->force_L, fkq))
      return fail(FKPM_E_UNSUP, "geometry lanes=%d kq=%d not compiled for block %d", h->force_L,
                  fkq, h->b);
    wmax = std::max(wmax, h->force_L * fkq);
  }
  h->wmax = wmax;
  h->qp = ((qp0 + wmax - 1) / wmax) * wmax;
  h->sp = h->qp * ne;
  h->geo_b = h->b;
  bool cached = false;
  std::array<KCfg, NKIND> cc;
  if (h->force_L == 0) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(tune_key(h));
    if (it != g_tune_cache.end()) {
      cc = it->second;
      cached = true;
    }
  }
  for (int kind = 0; kind < NKIND; kind++) {
    KCfg c;
    if (h->force_L > 0) {
      c.L = h->force_L;
      c.KQ = h->force_KQ > 0 ? h->force_KQ : 1;
      c.maxwg = h->max_wg;
    } else if (cached) {
      c = cc[kind];
    } else {
      c = default_cfg(h, kind);
      if (h->max_wg > 0) c.maxwg = h->max_wg;
    }
    CHK(set_kind(h, kind, c));
  }
  h->tuned = cached;
  return 0;
}

template <class TH, class S, bool CX>
int pack_R_t(fkpm_handle* h) {
  dim3 grid((h->n + 63) / 64, (h->sp + 15) / 16);
  CHK(launch(&fkd::k_pack_R<TH, S, CX>, grid, dim3(fkd::WG), h->stream, h->rraw.as<const TH>(),
             h->R.as<S>(), h->n, h->s, h->sp));
  // compact form when every row holds at most one nonzero
  CHK(h->rcol.ensure(sizeof(int) * (size_t)h->n));
  CHK(h->rval.ensure(sizeof(S) * (size_t)h->n * (CX ? 2 : 1)));
  CHK(h->rflag.ensure(16));
  CHK(launch(&fkd::k_zero, dim3(1), dim3(64), h->stream, (double*)h->rflag.p, (size_t)2));
  CHK(launch(&fkd::k_compact_R<TH, S, CX>, dim3((h->n + fkd::WG - 1) / fkd::WG), dim3(fkd::WG),
             h->stream, h->rraw.as<const TH>(), h->n, h->s, h->rcol.as<int>(), h->rval.as<S>(),
             h->rflag.as<int>()));
  int flag = 1;
  HIPCHK(hipMemcpyAsync(&flag, h->rflag.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->r_compact = (flag == 0);
  return 0;
}

template <class S>
int pack_R_s(fkpm_handle* h) {
  switch (h->dtype) {
    case FKPM_F32: return pack_R_t<float, S, false>(h);
    case FKPM_F64: return pack_R_t<double, S, false>(h);
    case FKPM_C32: return pack_R_t<float, S, true>(h);
    default: return pack_R_t<double, S, true>(h);
  }
}

template <class S>
fkd::RSource<S> rsource(fkpm_handle* h, double cr) {
  fkd::RSource<S> r = no_rsource<S>();
  r.cr = cr;
  if (h->r_compact) {
    r.rcol = h->rcol.as<const int>();
    r.rval = h->rval.as<const S>();
  } else {
    r.rd = h->R.as<const S>();
  }
  return r;
}

int tune(fkpm_handle* h);

// Make geometry, vectors and packed R consistent with the current H and R.
int ensure_ready(fkpm_handle* h) {
  if (!h->h_set) return fail(FKPM_E_STATE, "set_H has not been called");
  if (!h->r_set) return fail(FKPM_E_STATE, "R has not been set (set_R / transfer_R)");
  if (h->n != h->nb * h->b)
    return fail(FKPM_E_ARG, "R has %d rows but H has %d = %d blocks x %d", h->n, h->nb * h->b,
                h->nb, h->b);
  bool stale = h->r_dirty || h->geo_b != h->b;
  // backstop: a row-resident matrix-core kernel on a pattern with longer rows
  // would read only the first kRowResidentMax blocks of each row
  for (int k = 0; k < NKIND && !stale; k++)
    stale = h->kc[k].mfma == 2 && (!mfma_rr_avail(h, k) || (h->kc[k].dm == 7 && h->max_deg > 7));
  if (!stale) return 0;
  CHK(choose_geometry(h));
  const size_t vb = h->vbytes();
  CHK(h->R.ensure(vb));
  for (int k = 0; k < 3; k++) CHK(h->A[k].ensure(vb));
  if (h->single)
    CHK(pack_R_s<float>(h));
  else
    CHK(pack_R_s<double>(h));
  // A-buffers: every element is written by the recursion before it is read.
  h->r_dirty = false;
  h->alpha_M = 0;
  h->tab_K = 0;
  if (h->autotune_on_use && !h->tuned && h->force_L == 0) CHK(tune(h));
  return 0;
}

int ensure_partials(fkpm_handle* h, int nsteps) {
  const int nwg = std::max(h->kc[K_MOM].nwg, h->kc[K_MOMT].nwg);
  CHK(h->partial.ensure((size_t)nsteps * nwg * 2 * sizeof(double)));
  CHK(h->psum.ensure((size_t)nsteps * 2 * sizeof(double)));
  return 0;
}

int zero_doubles(fkpm_handle* h, void* p, size_t n) {
  return launch(&fkd::k_zero, dim3(grid_for(n)), dim3(256), h->stream, (double*)p, n);
}

int timing_begin(fkpm_handle* h) {
  if (h->timing) HIPCHK(hipEventRecord(h->ev0, h->stream));
  return 0;
}
int timing_end(fkpm_handle* h, double* ms, int* cnt, int launches) {
  if (!h->timing) return 0;
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipEventSynchronize(h->ev1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, h->ev0, h->ev1));
  *ms = t;
  *cnt = launches;
  return 0;
}

// ---------------------------------------------------------------------------
// Autotuner: times every compiled geometry of pass width wmax or wmax/2, three
// workgroup caps and the nontemporal choice, per kernel kind, on the current
// H and R; keeps the fastest and caches it for this problem shape.
// ---------------------------------------------------------------------------
template <class S>
int tune_t(fkpm_handle* h) {
  h->tab_K = 0;  // the table may serve as tuning slots below
  h->alpha_M = 0;
  const int HS = h->cx ? 2 : 1;
  const size_t vb = h->vbytes();
  CHK(h->Bv[0].ensure(vb));
  CHK(h->xi.ensure(h->accbytes()));
  CHK(h->tune_dacc.ensure((size_t)(h->nnzb > 0 ? h->nnzb : 1) * h->b * h->b * HS * sizeof(double)));
  CHK(h->partial.ensure((size_t)(h->nb + 3) / 4 * 2 * sizeof(double) + 64));
  const int* rp = h->rp.as<const int>();
  const int* ci = h->ci.as<const int>();
  const S* hv = h->hv.as<const S>();
  const S* R = h->R.as<const S>();
  S* A0 = h->A[0].as<S>();
  S* A1 = h->A[1].as<S>();
  S* A2 = h->A[2].as<S>();
  S* B0 = h->Bv[0].as<S>();
  CHK(h->Bv[1].ensure(vb));
  CHK(h->Bv[2].ensure(vb));
  S* B1 = h->Bv[1].as<S>();
  S* B2 = h->Bv[2].as<S>();
  // finite inputs for every candidate: every ring buffer = Hs R
  {
    const KCfg& c = h->kc[K_MOM];
    const auto& ks = h->ks<S>(K_MOM);
    S* outs[] = {A0, A1, A2, B0, B1, B2};
    for (S* o : outs)
      CHK(launch(mom_fn(h, ks, c), dim3(c.nwg), dim3(fkd::WG), h->stream, rp, ci, hv, R, (const S*)nullptr,
                 o, h->nb, h->qp, c.chunk, 1.0, 0.0, (double*)nullptr, 0, no_rsource<S>()));
    CHK(zero_doubles(h, h->tune_dacc.p, (size_t)h->nnzb * h->b * h->b * HS));
  }
  S* ring[3] = {A0, A1, A2};
  S* bring[3] = {B0, B1, B2};
  const bool log = getenv("FKPM_TUNE_LOG") != nullptr;
  int step = 0;
  fkd::GeomChoice gs[64];
  const int ng = fkd::geometries_for_block(h->b, gs, 64);
  static const int caps[] = {1024, 4096, 16384};
  static const int nts[] = {0, fkd::NT_LOAD | fkd::NT_STORE};
  static const int sweeps[] = {0, 1, 2, 8};
  // The table kinds are timed on ns slots (at least 512 MB, more than the
  // last-level cache holds) walked as the recursions walk the table: a fresh
  // slot per forward step, the kept vectors backwards.  Without a table in
  // prospect they take the ring kinds' choice.
  int ns = 0;
  int tstep = 0;
  auto slot = [&](int k) { return (S*)((char*)h->atab.p + (size_t)k * vb); };
  const int kinds[] = {K_MOM, K_XI, K_AD, K_MOMT, K_ADS};
  for (int kind : kinds) {
    if (kind == K_MOMT) {
      if (table_wanted(h) && !h->kc[K_AD].mfma) {
        const size_t want = ((size_t)512 << 20) / vb + 1;
        const int n = (int)std::min<size_t>(64, std::max<size_t>(8, want));
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        if ((size_t)n * vb <= free_b / 4 + h->atab.bytes) {
          CHK(h->atab.ensure((size_t)n * vb));
          ns = n;
          const KCfg& c = h->kc[K_MOM];
          const auto& ks = h->ks<S>(K_MOM);
          for (int k = 0; k < ns; k++)  // every slot = Hs R
            CHK(launch(mom_fn(h, ks, c), dim3(c.nwg), dim3(fkd::WG), h->stream, rp, ci, hv, R,
                       (const S*)nullptr, slot(k), h->nb, h->qp, c.chunk, 1.0, 0.0,
                       (double*)nullptr, 0, no_rsource<S>()));
        }
      }
      if (ns == 0) {
        CHK(set_kind(h, K_MOMT, h->kc[K_MOM]));
        // no table while autodiff runs on the matrix cores
        CHK(set_kind(h, K_ADS, h->kc[K_AD].mfma ? default_cfg(h, K_ADS) : h->kc[K_AD]));
        break;
      }
    }
    KCfg best = h->kc[kind];
    float best_ms = 1e30f;
    std::vector<KCfg> cands;
    for (int g = 0; g < ng; g++) {
      const int w = gs[g].L * gs[g].KQ;
      if (w > h->wmax || 2 * w < h->wmax) continue;
      // the column passes as the outer loop (NT_SLICE) when a row takes
      // several passes of this width
      const int nsl = h->qp / w >= 2 ? 2 : 1;
      for (int cap : caps)
        for (int nt : nts)
          for (int sl = 0; sl < nsl; sl++) {
            KCfg c;
            c.L = gs[g].L;
            c.KQ = gs[g].KQ;
            c.maxwg = cap;
            c.nt = nt | (sl ? fkd::NT_SLICE : 0);
            for (int sw : sweeps) {
              c.sweep = sw;
              cands.push_back(c);
            }
            if (h->has_order && kind != K_OUT) {  // the lattice row order, one range per XCD
              c.nt |= fkd::NT_ORDER;
              c.sweep = 0;
              cands.push_back(c);
            }
          }
    }
    if (mfma_avail(h, kind))
      for (int v = 1; v <= (mfma_rr_avail(h, kind) ? 2 : 1); v++)
        for (int cap : caps)
          for (int sw : sweeps) {
            KCfg c;
            c.mfma = v;
            c.maxwg = cap;
            c.sweep = sw;
            cands.push_back(c);
          }
    // every candidate timed over 6 launches, then the closest few again over
    // 24, so that one noisy short timing does not decide the kernel for every
    // later solve of this shape
    auto time_cfg = [&](const KCfg& c, int reps, float* us) -> int {
          CHK(set_kind(h, kind, c));
          const KCfg& k = h->kc[kind];
          const auto& ks = h->ks<S>(kind);
          // steps rotate their buffers as the recursions do (one continuous
          // recurrence over all candidates keeps the values bounded)
          auto run = [&]() -> int {
            if (kind == K_MOMT) {  // alpha_{t+2} = 2 Hs alpha_{t+1} - alpha_t into a fresh slot
              const int t = tstep++;
              return launch(mom_fn(h, ks, k), dim3(k.nwg), dim3(fkd::WG), h->stream, rp_for(h, k), ci, hv,
                            (const S*)slot((t + 1) % ns), (const S*)slot(t % ns),
                            slot((t + 2) % ns), h->nb, h->qp, k.chunk, 2.0, -1.0,
                            h->partial.as<double>(), k.nt, no_rsource<S>());
            }
            if (kind == K_ADS) {  // backward over slots ns-2 .. 1, beta in the ring
              const int t = tstep++;
              const int m = ns - 2 - t % (ns - 2);
              return launch(ks.ad_stored, dim3(k.nwg), dim3(fkd::WG), h->stream, rp_for(h, k), ci,
                            hv, (const S*)slot(m), (const S*)slot(m + 1),
                            (const S*)bring[(t + 1) % 3], (const S*)bring[t % 3],
                            (const S*)slot(m - 1), bring[(t + 2) % 3], h->tune_dacc.as<double>(),
                            h->nb, h->qp, k.chunk, 0.1, 0.1, 0.1, 1.0, 0.0, 2.0, k.nt);
            }
            const int t = step++;
            S* x = ring[(t + 1) % 3];
            S* z = ring[t % 3];
            S* y = ring[(t + 2) % 3];
            // recurrences keep |x| bounded: y = 2 Hs x - z stays within the spectrum scale.
            // The forward kernel also runs stoch_matrix's Clenshaw steps (with
            // the source term c_k R, without the moment dots): every other
            // launch here is one, so both forms weigh in the choice
            if (kind == K_MOM) {
              const bool cl = t & 1;
              return launch(mom_fn(h, ks, k), dim3(k.nwg), dim3(fkd::WG), h->stream, rp_for(h, k), ci, hv,
                            (const S*)x, (const S*)z, y, h->nb, h->qp, k.chunk, 2.0, -1.0,
                            cl ? (double*)nullptr : h->partial.as<double>(), k.nt,
                            cl ? rsource<S>(h, 0.25) : no_rsource<S>());
            }
            if (kind == K_XI)
              return launch(ks.xi, dim3(k.nwg), dim3(fkd::WG), h->stream, rp_for(h, k), ci, hv, (const S*)x,
                            (const S*)z, y, h->xi.as<double>(), h->nb, h->qp, k.chunk, 2.0, -1.0,
                            0.0, 0.5, t == 0 ? 1 : 0, k.nt);
            // backward step: reads alpha_m, alpha_{m+1}, beta_{m+1}, beta_{m+2};
            // writes alpha_{m-1}, beta_m into the third buffers
            return launch(adf_fn(h, ks, k), dim3(k.nwg), dim3(fkd::WG), h->stream, rp_for(h, k), ci, hv,
                          (const S*)x, (const S*)z, (const S*)bring[(t + 1) % 3],
                          (const S*)bring[t % 3], y, bring[(t + 2) % 3], h->tune_dacc.as<double>(),
                          h->nb, h->qp, k.chunk, 0.1, 0.1, 0.1, 1.0, 0.0, 2.0, k.nt);
          };
          CHK(run());
          HIPCHK(hipEventRecord(h->ev0, h->stream));
          for (int r = 0; r < reps; r++) CHK(run());
          HIPCHK(hipEventRecord(h->ev1, h->stream));
          HIPCHK(hipEventSynchronize(h->ev1));
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
          *us = 1e3f * ms / reps;
          if (log)
            fprintf(stderr,
                    "[fkpm tune] kind=%d L=%d KQ=%d mfma=%d maxwg=%d nwg=%d nt=%d sweep=%d: %.2f us%s\n",
                    kind, k.L, k.KQ, k.mfma, k.maxwg, k.nwg, k.nt, k.sweep, *us,
                    reps > 6 ? " (final)" : "");
          return 0;
    };
    std::vector<std::pair<float, KCfg>> timed;
    for (const KCfg& c : cands) {
      float us = 0.f;
      CHK(time_cfg(c, 6, &us));
      timed.push_back({us, h->kc[kind]});
    }
    std::stable_sort(timed.begin(), timed.end(),
                     [](const std::pair<float, KCfg>& a, const std::pair<float, KCfg>& b) {
                       return a.first < b.first;
                     });
    const size_t nfinal = std::min<size_t>(4, timed.size());
    for (size_t f = 0; f < nfinal; f++) {
      float us = 0.f;
      CHK(time_cfg(timed[f].second, 24, &us));
      if (us < best_ms) {
        best_ms = us;
        best = h->kc[kind];
      }
    }
    CHK(set_kind(h, kind, best));
  }
  // outer products: once per call, the autodiff geometry (a vector-unit
  // geometry of full pass width when autodiff runs on the matrix cores)
  KCfg outc = h->kc[K_AD].mfma ? default_cfg(h, K_OUT) : h->kc[K_AD];
  outc.nt = 0;
  CHK(set_kind(h, K_OUT, outc));
  h->alpha_M = 0;
  h->tab_K = 0;  // the table slots held tuning data
  h->tuned = true;
  std::array<KCfg, NKIND> cc;
  for (int k = 0; k < NKIND; k++) cc[k] = h->kc[k];
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[tune_key(h)] = cc;
  return 0;
}

int tune(fkpm_handle* h) { return h->single ? tune_t<float>(h) : tune_t<double>(h); }

// ---------------------------------------------------------------------------
// moments (engine_cpu.cpp:63-84; engine_cusparse.cpp:755-791)
// ---------------------------------------------------------------------------
template <class S>
int moments_t(fkpm_handle* h, int M, double* mu) {
  const int nsteps = M / 2;
  CHK(ensure_partials(h, nsteps));
  // keep every alpha_k for autodiff_matrix (no backward recomputation of
  // alpha) when wanted (table_wanted), the table fits half the free device
  // memory (or the byte count in FKPM_ALPHA_TABLE_BYTES) and the autodiff
  // step runs on the vector units.
  {
    const size_t need = (size_t)nsteps * h->vbytes();
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    // default: half of what is free (the table already held counts as free),
    // never more than a quarter of the device, so that other engines and a
    // batch arena on the same device keep room whatever the order of calls
    size_t budget = std::min(free_b / 2 + h->atab.bytes, total_b / 4);
    if (h->table_cap_mb >= 0) budget = (size_t)h->table_cap_mb << 20;
    if (const char* e = getenv("FKPM_ALPHA_TABLE_BYTES")) budget = (size_t)strtoull(e, nullptr, 10);
    h->a_stored = false;
    if (table_wanted(h) && !h->kc[K_AD].mfma && need <= budget) {
      h->tab_K = 0;  // rewritten below (and lost if the table grows)
      CHK(h->atab.ensure(need));
      h->a_stored = true;
    }
  }
  const int kind = h->a_stored ? K_MOMT : K_MOM;
  const KCfg& kc = h->kc[kind];
  const auto& ks = h->ks<S>(kind);
  const dim3 grid(kc.nwg), block(fkd::WG);
  CHK(timing_begin(h));
  {
    LoopRecorder rec(h);
    for (int m = 0; m < nsteps; m++) {
      // alpha_{m+1} = cx Hs alpha_m + cz alpha_{m-1}
      const S* x = (const S*)h->vec_a(m);
      const S* z = m == 0 ? nullptr : (const S*)h->vec_a(m - 1);
      S* y = (S*)h->vec_a(m + 1);
      CHK(launch(mom_fn(h, ks, kc), grid, block, h->stream, rp_for(h, kc),
                 h->ci.as<const int>(), h->hv.as<const S>(), x, z, y, h->nb, h->qp, kc.chunk,
                 m == 0 ? 1.0 : 2.0, m == 0 ? 0.0 : -1.0,
                 h->partial.as<double>() + (size_t)m * kc.nwg * 2, kc.nt, no_rsource<S>()));
    }
    CHK(launch(&fkd::k_reduce_partials, dim3(nsteps), block, h->stream,
               (const double*)h->partial.as<double>(), kc.nwg, h->psum.as<double>()));
    CHK(rec.run());
  }
  CHK(timing_end(h, &h->ms_mom, &h->n_mom, nsteps));
  if (h->a_stored) h->tab_K = nsteps;
  std::vector<double> ps((size_t)nsteps * 2);
  HIPCHK(hipMemcpyAsync(ps.data(), h->psum.p, ps.size() * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  mu[0] = ps[0];
  mu[1] = ps[1];
  for (int m = 1; m < nsteps; m++) {
    mu[2 * m] = 2 * ps[2 * m] - mu[0];
    mu[2 * m + 1] = 2 * ps[2 * m + 1] - mu[1];
  }
  h->alpha_M = M;
  return 0;
}

// D to the host in the engine type.
int copy_D_out(fkpm_handle* h, void* D_val) {
  const size_t cnt = (size_t)h->nnzbD * h->b * h->b * (h->cx ? 2 : 1);
  if (!dtype_single(h->dtype)) {
    HIPCHK(hipMemcpyAsync(D_val, h->dout.p, cnt * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  } else {
    std::vector<double> tmp(cnt);
    HIPCHK(hipMemcpyAsync(tmp.data(), h->dout.p, cnt * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    float* o = static_cast<float*>(D_val);
    for (size_t k = 0; k < cnt; k++) o[k] = (float)tmp[k];
  }
  return 0;
}

int check_D_ready(fkpm_handle* h) {
  if (!h->d_set) return fail(FKPM_E_STATE, "D pattern has not been set");
  if (h->nbD != h->nb)
    return fail(FKPM_E_ARG, "D has %d block rows but H has %d", h->nbD, h->nb);
  const size_t dbytes = (size_t)h->nnzbD * h->b * h->b * (h->cx ? 2 : 1) * sizeof(double);
  CHK(h->dacc.ensure(dbytes > 0 ? dbytes : 16));
  CHK(h->dout.ensure(dbytes > 0 ? dbytes : 16));
  h->d_fused = (h->hd_rp == h->h_rp && h->hd_ci == h->h_ci);
  return 0;
}

// ---------------------------------------------------------------------------
// stoch_matrix (engine_cpu.cpp:229-253; engine_cusparse.cpp:910-953)
// ---------------------------------------------------------------------------

// xi = sum_{m<M} c_m T_m(Hs) R by the Clenshaw recurrence
//   b_{M-1} = c_{M-1} R, b_M = 0,
//   b_k = c_k R + 2 Hs b_{k+1} - b_{k+2}   (k = M-2 .. 1),
//   xi = c_0 R + Hs b_1 - b_2,
// one SpMM pass per step reading two vectors (no FP64 accumulator traffic);
// b_{M-1} stays virtual (a multiple of R).  Returns xi's buffer.
template <class S>
int stoch_clenshaw(fkpm_handle* h, int M, const double* c, const S** xi_out) {
  const KCfg& kc = h->kc[K_MOM];
  const auto& ks = h->ks<S>(K_MOM);
  const dim3 grid(kc.nwg), block(fkd::WG);
  const S* R = (const S*)h->R.p;
  struct Vb {
    const S* p;
    double s;
  };
  Vb b1{R, c[M - 1]}, b2{nullptr, 0.0};
  int slot = 0, launches = 0;
  auto step = [&](double xfac, double cr, S* y) -> int {
    return launch(mom_fn(h, ks, kc), grid, block, h->stream, rp_for(h, kc), h->ci.as<const int>(),
                  h->hv.as<const S>(), b1.p, b2.p, y, h->nb, h->qp, kc.chunk, xfac * b1.s,
                  b2.p ? -b2.s : 0.0, (double*)nullptr, kc.nt, rsource<S>(h, cr));
  };
  for (int k = M - 2; k >= 1; k--) {
    S* y = (S*)h->A[slot].p;
    slot = (slot + 1) % 3;
    CHK(step(2.0, c[k], y));
    launches++;
    b2 = b1;
    b1 = Vb{y, 1.0};
  }
  S* y = (S*)h->A[slot].p;
  CHK(step(1.0, c[0], y));
  *xi_out = y;
  return launches + 1;
}

// xi from the kept forward vectors alpha_0 .. alpha_K (K = ceil(M/2), coefficients
// padded with zeros to 2K): by T_{K+j} = 2 T_K T_j - T_{K-j},
//   xi = sum_{m<2K} c_m T_m(Hs) R = v + 2 T_K(Hs) u,
//   v = sum_{i=0..K} ([i<K] c_i - [i>=1] c_{2K-i}) alpha_i,
//   u = sum_{i<K} c_{K+i} alpha_i,
// one streaming pass over the table (k_table_sums) and K Chebyshev steps
// from u, the last adding v as its source term: about 4K + 4 vector
// transfers instead of the 3(M - 1) of the Clenshaw recurrence.  Returns
// xi's buffer.
template <class S>
int stoch_table(fkpm_handle* h, int M, const double* c, const S** xi_out) {
  const int K = (M + 1) / 2;
  const KCfg& kc = h->kc[K_MOM];
  const auto& ks = h->ks<S>(K_MOM);
  const dim3 grid(kc.nwg), block(fkd::WG);
  const size_t vb = h->vbytes();
  auto cm = [&](int m) { return m < M ? c[m] : 0.0; };
  std::vector<double>& w = h->tabc_host;
  w.assign(2 * (size_t)(K + 1), 0.0);
  for (int i = 0; i <= K; i++) {
    w[i] = (i < K ? cm(i) : 0.0) - (i >= 1 ? cm(2 * K - i) : 0.0);  // d_i
    w[K + 1 + i] = i < K ? cm(K + i) : 0.0;                         // e_i
  }
  CHK(h->tabc.ensure(w.size() * sizeof(double)));
  HIPCHK(hipMemcpyAsync(h->tabc.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice,
                        h->stream));
  CHK(h->Bv[0].ensure(vb));
  S* u = (S*)h->A[0].p;
  S* v = (S*)h->Bv[0].p;
  const size_t n = vb / sizeof(S);
  CHK(launch(&fkd::k_table_sums<S>, dim3(4096), block, h->stream, (const S*)h->R.p,
             (const S*)h->atab.p, n, K, (const double*)h->tabc.as<double>(),
             (const double*)h->tabc.as<double>() + K + 1, u, v, n));
  // y_0 = u, y_1 = Hs u, y_{k+1} = 2 Hs y_k - y_{k-1}; the last step writes
  // 2 y_K + v
  fkd::RSource<S> addv = no_rsource<S>();
  addv.rd = v;
  addv.cr = 1.0;
  const S* x = u;
  const S* z = nullptr;
  int slot = 1;
  for (int k = 0; k < K; k++) {
    S* y = (S*)h->A[slot].p;
    slot = (slot + 1) % 3;
    const bool last = (k == K - 1);
    const double cx = (k == 0 ? 1.0 : 2.0) * (last ? 2.0 : 1.0);
    const double cz = k == 0 ? 0.0 : (last ? -2.0 : -1.0);
    CHK(launch(mom_fn(h, ks, kc), grid, block, h->stream, rp_for(h, kc),
               h->ci.as<const int>(), h->hv.as<const S>(), x, z, y, h->nb, h->qp, kc.chunk, cx,
               cz, (double*)nullptr, kc.nt, last ? addv : no_rsource<S>()));
    z = x;
    x = y;
  }
  *xi_out = x;
  return K + 1;
}

template <class S>
int stoch_t(fkpm_handle* h, int M, const double* c) {
  h->a_stored = false;  // the forward method's alpha_k go to the ring
  const dim3 block(fkd::WG);
  const int* rp = h->rp.as<const int>();
  const int* ci = h->ci.as<const int>();
  const S* hv = h->hv.as<const S>();
  const size_t dcount = (size_t)h->nnzbD * h->b * h->b * (h->cx ? 2 : 1);
  const KCfg& ko = h->kc[K_OUT];
  const auto& kso = h->ks<S>(K_OUT);
  CHK(timing_begin(h));
  const bool table = h->stoch_method == FKPM_STOCH_AUTO && h->tab_K >= (M + 1) / 2;
  h->stoch_used = table ? FKPM_STOCH_TABLE
                        : h->stoch_method == FKPM_STOCH_FORWARD ? FKPM_STOCH_FORWARD
                                                                : FKPM_STOCH_CLENSHAW;
  if (h->stoch_used != FKPM_STOCH_FORWARD) {
    const S* xi = nullptr;
    int launches = 0;
    {
      LoopRecorder rec(h);
      launches = table ? stoch_table<S>(h, M, c, &xi) : stoch_clenshaw<S>(h, M, c, &xi);
      if (launches < 0) return launches;
      CHK(rec.run());
    }
    CHK(timing_end(h, &h->ms_stoch, &h->n_stoch, launches));
    CHK(zero_doubles(h, h->dacc.p, dcount));
    // D = (1/2) R xi^+ + (1/2) xi R^+
    CHK(launch(kso.outer_ss, dim3(ko.nwg), block, h->stream, h->drp.as<const int>(),
               h->dci.as<const int>(), (const S*)h->R.p, xi, h->dacc.as<double>(), h->nb, h->qp,
               ko.chunk, 0.5));
    CHK(launch(kso.outer_ss, dim3(ko.nwg), block, h->stream, h->drp.as<const int>(),
               h->dci.as<const int>(), xi, (const S*)h->R.p, h->dacc.as<double>(), h->nb, h->qp,
               ko.chunk, 0.5));
  } else {
    // the reference's order: forward recursion, xi accumulated in FP64
    CHK(h->xi.ensure(h->accbytes()));
    const KCfg& kx = h->kc[K_XI];
    const auto& ksx = h->ks<S>(K_XI);
    const dim3 grid(kx.nwg);
    {
      LoopRecorder rec(h);
      // alpha_1 = Hs R, xi = c0 R + c1 alpha_1
      CHK(launch(ksx.xi, grid, block, h->stream, rp_for(h, kx), ci, hv, (const S*)h->vec_a(0),
                 (const S*)nullptr, (S*)h->vec_a(1), h->xi.as<double>(), h->nb, h->qp, kx.chunk,
                 1.0, 0.0, c[0], c[1], 1, kx.nt));
      for (int m = 2; m < M; m++) {
        // alpha_m = 2 Hs alpha_{m-1} - alpha_{m-2}; xi += c_m alpha_m
        CHK(launch(ksx.xi, grid, block, h->stream, rp_for(h, kx), ci, hv,
                   (const S*)h->vec_a(m - 1),
                   (const S*)h->vec_a(m - 2), (S*)h->vec_a(m), h->xi.as<double>(), h->nb, h->qp,
                   kx.chunk, 2.0, -1.0, 0.0, c[m], 0, kx.nt));
      }
      CHK(rec.run());
    }
    CHK(timing_end(h, &h->ms_stoch, &h->n_stoch, M - 1));
    CHK(zero_doubles(h, h->dacc.p, dcount));
    CHK(launch(kso.outer_sd, dim3(ko.nwg), block, h->stream, h->drp.as<const int>(),
               h->dci.as<const int>(), (const S*)h->vec_a(0), (const double*)h->xi.as<double>(),
               h->dacc.as<double>(), h->nb, h->qp, ko.chunk, 0.5));
    CHK(launch(kso.outer_ds, dim3(ko.nwg), block, h->stream, h->drp.as<const int>(),
               h->dci.as<const int>(), (const double*)h->xi.as<double>(), (const S*)h->vec_a(0),
               h->dacc.as<double>(), h->nb, h->qp, ko.chunk, 0.5));
  }
  HIPCHK(hipMemcpyAsync(h->dout.p, h->dacc.p, dcount * sizeof(double), hipMemcpyDeviceToDevice,
                        h->stream));
  h->alpha_M = 0;  // engine_cpu.cpp:250-252
  return 0;
}

// ---------------------------------------------------------------------------
// autodiff_matrix (engine_cpu.cpp:255-304; engine_cusparse.cpp:955-1039)
// ---------------------------------------------------------------------------
template <class S>
int autodiff_t(fkpm_handle* h, int M, const double* c) {
  const dim3 block(fkd::WG);
  const KCfg& ka = h->kc[K_AD];
  const KCfg& kt = h->kc[K_ADS];
  const KCfg& ko = h->kc[K_OUT];
  const auto& ksa = h->ks<S>(K_AD);
  const auto& kst = h->ks<S>(K_ADS);
  const auto& kso = h->ks<S>(K_OUT);
  const int* rp = h->rp.as<const int>();
  const int* ci = h->ci.as<const int>();
  const S* hv = h->hv.as<const S>();
  const int* drp = h->drp.as<const int>();
  const int* dci = h->dci.as<const int>();
  double* dacc = h->dacc.as<double>();
  const size_t dcount = (size_t)h->nnzbD * h->b * h->b * (h->cx ? 2 : 1);

  double diag = c[1];
  for (int m = 1; m < M / 2; m++) diag -= c[2 * m + 1];
  CHK(zero_doubles(h, dacc, dcount));
  // D = diag R R^+
  CHK(launch(kso.outer_ss, dim3(ko.nwg), block, h->stream, drp, dci, (const S*)h->vec_a(0),
             (const S*)h->vec_a(0), dacc, h->nb, h->qp, ko.chunk, diag));
  int launches = 0;
  if (M > 2)
    for (int k = 0; k < 3; k++) CHK(h->Bv[k].ensure(h->vbytes()));
  CHK(timing_begin(h));
  LoopRecorder rec(h);
  if (M > 2) {
    const double sb = 2 * c[M - 1];  // beta_{M/2} = 2 c_{M-1} alpha_{M/2-1}
    const bool stored = h->a_stored && !ka.mfma;
    for (int m = M / 2 - 1; m >= 1; m--) {
      // alpha_m, alpha_{m+1} and alpha_{m-1}: from the table, or the ring with
      // alpha_{m-1} recomputed by the step
      const S* am = (const S*)(stored ? h->vec_a(m) : h->A[m % 3].p);
      const S* am1 = (const S*)(stored ? h->vec_a(m + 1) : h->A[(m + 1) % 3].p);
      S* aout = stored ? (S*)h->vec_a(m - 1) : (S*)h->A[(m - 1) % 3].p;
      const bool first = (m == M / 2 - 1);
      const bool second = (m == M / 2 - 2);
      const S* bm1 = first ? nullptr : (const S*)h->Bv[(m + 1) % 3].p;
      const S* bm2 = (first || second) ? nullptr : (const S*)h->Bv[(m + 2) % 3].p;
      const double sb2 = second ? sb : 0.0;
      S* bout = (S*)h->Bv[m % 3].p;
      const double c4 = 4 * c[2 * m], c2 = 2 * c[2 * m + 1];
      const double w = m > 1 ? 2 * c[2 * m - 1] : 0.0;
      if (stored) {
        if (!h->d_fused)  // D += 2 alpha_m beta_{m+1}^+ on D's own pattern
          CHK(launch(kso.outer_ss, dim3(ko.nwg), block, h->stream, drp, dci, am, first ? am : bm1,
                     dacc, h->nb, h->qp, ko.chunk, first ? 2.0 * sb : 2.0));
        CHK(launch(h->d_fused ? kst.ad_stored : kst.ad_stored_plain, dim3(kt.nwg), block,
                   h->stream, rp_for(h, kt), ci, hv, am, am1, bm1, bm2, (const S*)aout, bout, dacc,
                   h->nb, h->qp, kt.chunk, c4, c2, w, sb, sb2, 2.0, kt.nt));
      } else if (h->d_fused) {
        CHK(launch(adf_fn(h, ksa, ka), dim3(ka.nwg), block, h->stream, rp_for(h, ka), ci, hv, am, am1,
                   bm1, bm2,
                   aout, bout, dacc, h->nb, h->qp, ka.chunk, c4, c2, w, sb, sb2, 2.0, ka.nt));
      } else {
        // D += 2 alpha_m beta_{m+1}^+ on D's own pattern, then the step.
        CHK(launch(kso.outer_ss, dim3(ko.nwg), block, h->stream, drp, dci, am, first ? am : bm1,
                   dacc, h->nb, h->qp, ko.chunk, first ? 2.0 * sb : 2.0));
        CHK(launch(adp_fn(ksa, ka), dim3(ka.nwg), block, h->stream, rp_for(h, ka), ci, hv, am, am1,
                   bm1, bm2,
                   aout, bout, dacc, h->nb, h->qp, ka.chunk, c4, c2, w, sb, sb2, 2.0, ka.nt));
      }
      launches++;
    }
    // D += alpha_0 beta_1^+ (alpha_0 = R kept, or recomputed into A[0])
    CHK(launch(kso.outer_ss, dim3(ko.nwg), block, h->stream, drp, dci,
               (const S*)(stored ? h->vec_a(0) : h->A[0].p), (const S*)h->Bv[1].p, dacc, h->nb,
               h->qp, ko.chunk, 1.0));
  }
  CHK(rec.run());
  CHK(timing_end(h, &h->ms_ad, &h->n_ad, launches));
  // symmetrize + scale(1/mag)
  const size_t dcnt = (size_t)h->nnzbD * h->b * h->b;
  if (h->cx)
    CHK(launch(&fkd::k_symmetrize_scale<true>, dim3(grid_for(dcnt)), dim3(256), h->stream,
               (const double*)dacc, h->dout.as<double>(), h->tpos.as<const int>(), h->nnzbD, h->b,
               1.0 / h->mag));
  else
    CHK(launch(&fkd::k_symmetrize_scale<false>, dim3(grid_for(dcnt)), dim3(256), h->stream,
               (const double*)dacc, h->dout.as<double>(), h->tpos.as<const int>(), h->nnzbD, h->b,
               1.0 / h->mag));
  h->alpha_M = 0;  // engine_cpu.cpp:301-303
  return 0;
}

// ---------------------------------------------------------------------------
// Two-current moments (Engine_CPU::moments2_v1, engine_cpu.cpp:86-160;
// engine_cusparse.cpp:793-902):
//   mu[m1][m2] = sum_k <alpha_m1(k), j2 atild_m2(k)>,
//   alpha_m(k) = T_m(Hs) r_k,  atild_m(k) = T_m(Hs) j1 r_k.
// For a chunk of probe columns all M alpha_m and all M j2 atild_m are kept in
// HBM as the columns of two K x M matrices (K = n * padded columns), built by
// the step kernels (j1, j2 on their own block patterns), and the M x M block
// of dots A^H B is formed on the matrix cores (kpm_dots.h), accumulated in
// FP64 over chunks.
// ---------------------------------------------------------------------------

// The kernel set and geometry for a pass width of W packets (a compiled
// (L, KQ) with L * KQ == W, widest lane groups first).
template <class S>
bool kset_for_width(fkpm_handle* h, int W, fkd::KSet<S>* ks, int* L, int* KQ);
template <>
bool kset_for_width<double>(fkpm_handle* h, int W, fkd::KSet<double>* ks, int* L, int* KQ) {
  fkd::GeomChoice gs[64];
  const int ng = fkd::geometries_for_block(h->b, gs, 64);
  for (int want = 64; want >= 1; want /= 2)
    for (int k = 0; k < ng; k++)
      if (gs[k].L == want && gs[k].L * gs[k].KQ == W) {
        *L = gs[k].L;
        *KQ = gs[k].KQ;
        return h->cx ? fkd::find_kset_dc(h->b, *L, *KQ, ks) : fkd::find_kset_dr(h->b, *L, *KQ, ks);
      }
  return false;
}
template <>
bool kset_for_width<float>(fkpm_handle* h, int W, fkd::KSet<float>* ks, int* L, int* KQ) {
  fkd::GeomChoice gs[64];
  const int ng = fkd::geometries_for_block(h->b, gs, 64);
  for (int want = 64; want >= 1; want /= 2)
    for (int k = 0; k < ng; k++)
      if (gs[k].L == want && gs[k].L * gs[k].KQ == W) {
        *L = gs[k].L;
        *KQ = gs[k].KQ;
        return h->cx ? fkd::find_kset_fc(h->b, *L, *KQ, ks) : fkd::find_kset_fr(h->b, *L, *KQ, ks);
      }
  return false;
}

// C (M x M, FP64, row-major (m1, m2)) = [C +] A^H B for K x M matrices A, B
// of the storage type (column m at A + m * ld): k_dot_block over slices of
// 512 k interleaved across about 2048 workgroups, then k_dot_reduce adding
// the grid rows' partial blocks in order (at most 1 GiB of them).
template <class S>
int dot_block_acc(fkpm_handle* h, const S* A, const S* B, long long K, long long ld, int M,
                  bool first) {
  const int HS = h->cx ? 2 : 1;
  const int nbm = (M + fkd::DOT_BM - 1) / fkd::DOT_BM;
  const int nblk = nbm * nbm;
  const long long kslice = 512;
  const size_t cnt = (size_t)M * M * HS;
  const long long nslices = (K + kslice - 1) / kslice;
  long long ny = std::max(1, 2048 / nblk);
  ny = std::min(ny, nslices);
  ny = std::min<long long>(ny, std::max<long long>(1, ((size_t)1 << 30) / (cnt * sizeof(double))));
  ny = std::min<long long>(ny, 65535);
  CHK(h->m2_part.ensure((size_t)ny * cnt * sizeof(double)));
  double* part = h->m2_part.as<double>();
  if (h->cx)
    CHK(launch(&fkd::k_dot_block<S, true>, dim3(nblk, (int)ny), dim3(fkd::WG), h->stream, A, B,
               ld, 0LL, K, kslice, M, part));
  else
    CHK(launch(&fkd::k_dot_block<S, false>, dim3(nblk, (int)ny), dim3(fkd::WG), h->stream, A, B,
               ld, 0LL, K, kslice, M, part));
  return launch(&fkd::k_dot_reduce, dim3(grid_for(cnt)), dim3(fkd::WG), h->stream,
                (const double*)part, (int)ny, cnt, h->m2_c.as<double>(), first ? 1 : 0);
}

template <class TH, class S, bool CX>
int upload_current(fkpm_handle* h, int which, int nnzb, const int* rp, const int* ci,
                   const void* val) {
  const int nb = h->nb, b = h->b;
  std::vector<int> rb(nnzb > 0 ? nnzb : 1);
  for (int i = 0; i < nb; i++)
    for (int p = rp[i]; p < rp[i + 1]; p++) rb[p] = i;
  DevBuf &drp = h->j_rp[which], &dci = h->j_ci[which], &drb = h->j_rb[which];
  DevBuf &draw = h->j_raw[which], &dval = h->j_val[which];
  const size_t cnt = (size_t)nnzb * b * b;
  CHK(drp.ensure(sizeof(int) * (nb + 1)));
  CHK(dci.ensure(sizeof(int) * (nnzb > 0 ? nnzb : 1)));
  CHK(drb.ensure(sizeof(int) * rb.size()));
  CHK(draw.ensure(cnt * sizeof(TH) * (CX ? 2 : 1) + 16));
  CHK(dval.ensure(cnt * sizeof(S) * (CX ? 2 : 1) + 16));
  HIPCHK(hipMemcpyAsync(drp.p, rp, sizeof(int) * (nb + 1), hipMemcpyHostToDevice, h->stream));
  if (nnzb > 0) {
    HIPCHK(hipMemcpyAsync(dci.p, ci, sizeof(int) * nnzb, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(draw.p, val, cnt * sizeof(TH) * (CX ? 2 : 1), hipMemcpyHostToDevice,
                          h->stream));
  }
  HIPCHK(hipMemcpyAsync(drb.p, rb.data(), sizeof(int) * rb.size(), hipMemcpyHostToDevice,
                        h->stream));
  if (cnt > 0)
    CHK(launch(&fkd::k_scale_H<TH, S, CX>, dim3(grid_for(cnt)), dim3(256), h->stream,
               draw.as<const TH>(), dval.as<S>(), drb.as<const int>(), dci.as<const int>(), nnzb,
               b, 1.0, 0.0));
  HIPCHK(hipStreamSynchronize(h->stream));  // host arrays (rb, caller's) go out of scope
  return 0;
}

template <class S>
int upload_current_s(fkpm_handle* h, int which, int nnzb, const int* rp, const int* ci,
                     const void* val) {
  switch (h->dtype) {
    case FKPM_F32: return upload_current<float, S, false>(h, which, nnzb, rp, ci, val);
    case FKPM_F64: return upload_current<double, S, false>(h, which, nnzb, rp, ci, val);
    case FKPM_C32: return upload_current<float, S, true>(h, which, nnzb, rp, ci, val);
    default: return upload_current<double, S, true>(h, which, nnzb, rp, ci, val);
  }
}

template <class TH, class S, bool CX>
int pack_R_cols(fkpm_handle* h, int k0, int sc, int spc, S* dst) {
  const TH* src = h->rraw.as<const TH>() + (size_t)k0 * h->n * (CX ? 2 : 1);
  const dim3 grid((h->n + 63) / 64, (spc + 15) / 16);
  return launch(&fkd::k_pack_R<TH, S, CX>, grid, dim3(fkd::WG), h->stream, src, dst, h->n, sc, spc);
}

template <class S>
int pack_R_cols_s(fkpm_handle* h, int k0, int sc, int spc, S* dst) {
  switch (h->dtype) {
    case FKPM_F32: return pack_R_cols<float, S, false>(h, k0, sc, spc, dst);
    case FKPM_F64: return pack_R_cols<double, S, false>(h, k0, sc, spc, dst);
    case FKPM_C32: return pack_R_cols<float, S, true>(h, k0, sc, spc, dst);
    default: return pack_R_cols<double, S, true>(h, k0, sc, spc, dst);
  }
}

template <class S>
int moments2_t(fkpm_handle* h, int M, int col_chunk, size_t budget, double* mu) {
  const int HS = h->cx ? 2 : 1;
  const int ne = h->ne();
  const size_t es = sizeof(S);
  // columns per chunk: the requested count (or all), halved until 2M + 3
  // vectors fit the budget
  auto padded = [&](int sc, int* W) {
    const int need = (sc + ne - 1) / ne;
    fkd::GeomChoice gs[64];
    const int ng = fkd::geometries_for_block(h->b, gs, 64);
    int best = 0, widest = 0;
    for (int k = 0; k < ng; k++) {
      const int w = gs[k].L * gs[k].KQ;
      widest = std::max(widest, w);
      if (w >= need && (best == 0 || w < best)) best = w;
    }
    if (best == 0) best = widest;
    *W = best;
    return ((need + best - 1) / best) * best * ne;  // padded columns
  };
  int sc = col_chunk > 0 ? std::min(col_chunk, h->s) : h->s;
  int W = 0, spc = padded(sc, &W);
  auto vec_bytes = [&](int spc_) { return (size_t)h->n * spc_ * HS * es; };
  while (sc > 1 && (size_t)(2 * M + 3) * vec_bytes(spc) > budget) {
    sc = (sc + 1) / 2;
    spc = padded(sc, &W);
  }
  if ((size_t)(2 * M + 3) * vec_bytes(spc) > budget)
    return fail(FKPM_E_ARG, "moments2: M = %d vectors of %zu bytes exceed the %zu-byte budget", M,
                vec_bytes(spc), budget);
  // the engine's own forward-step configuration (tuned or default) when it
  // has this pass width, else the widest lane groups of that width
  fkd::KSet<S> ks;
  KCfg c;
  const KCfg& km = h->kc[K_MOM];
  if (!km.mfma && km.L * km.KQ == W) {
    ks = h->ks<S>(K_MOM);
    c = km;
  } else {
    int L = 0, KQ = 0;
    if (!kset_for_width<S>(h, W, &ks, &L, &KQ))
      return fail(FKPM_E_UNSUP, "moments2: no kernel of width %d for block %d", W, h->b);
    c.L = L;
    c.KQ = KQ;
    c.maxwg = 4096;
    derive_grid(h->nb, c);
  }
  const int qpc = spc / ne;
  const size_t vb = vec_bytes(spc);
  // the columns of A and B are spaced by a vector plus 4352 bytes: with K a
  // power of two a plain spacing would put the same k of every column in the
  // same HBM channel
  const size_t vpad = 4352;
  CHK(h->m2_a.ensure((size_t)M * (vb + vpad)));
  CHK(h->m2_b.ensure((size_t)M * (vb + vpad)));
  CHK(h->m2_ring.ensure(3 * vb));
  CHK(h->m2_c.ensure((size_t)M * M * HS * sizeof(double)));
  const size_t vlen = vb / es;            // scalars per vector
  const size_t vs = (vb + vpad) / es;     // scalars between columns of A, B
  S* A = h->m2_a.as<S>();
  S* Bm = h->m2_b.as<S>();
  S* ring = h->m2_ring.as<S>();
  const int* rp = h->rp.as<const int>();
  const int* ci = h->ci.as<const int>();
  const S* hv = h->hv.as<const S>();
  const dim3 grid(c.nwg), block(fkd::WG);
  auto step = [&](const int* rp_, const int* ci_, const S* hv_, const S* x, const S* z, S* y,
                  double cx, double cz) {
    // the row order belongs to H's pattern; the currents' products sweep rows
    return launch(ks.mom, grid, block, h->stream, rp_, ci_, hv_, x, z, y, h->nb, qpc, c.chunk,
                  cx, cz, (double*)nullptr, c.nt & ~fkd::NT_ORDER, no_rsource<S>());
  };
  const int* j1rp = h->j_rp[0].as<const int>();
  const int* j1ci = h->j_ci[0].as<const int>();
  const S* j1v = h->j_val[0].as<const S>();
  const int* j2rp = h->j_rp[1].as<const int>();
  const int* j2ci = h->j_ci[1].as<const int>();
  const S* j2v = h->j_val[1].as<const S>();
  bool first = true;
  for (int k0 = 0; k0 < h->s; k0 += sc) {
    const int scc = std::min(sc, h->s - k0);
    CHK(pack_R_cols_s<S>(h, k0, scc, spc, A));
    // alpha_m = T_m(Hs) r into column m of A
    for (int m = 1; m < M; m++)
      CHK(step(rp, ci, hv, A + (m - 1) * vs, m >= 2 ? A + (m - 2) * vs : nullptr, A + m * vs,
               m == 1 ? 1.0 : 2.0, m == 1 ? 0.0 : -1.0));
    // atild_m = T_m(Hs) j1 r in a ring of three; column m of B = j2 atild_m
    CHK(step(j1rp, j1ci, j1v, A, nullptr, ring, 1.0, 0.0));
    for (int m = 0; m < M; m++) {
      S* t = ring + (m % 3) * vlen;
      if (m >= 1)
        CHK(step(rp, ci, hv, ring + ((m - 1) % 3) * vlen, m >= 2 ? ring + ((m - 2) % 3) * vlen : nullptr,
                 t, m == 1 ? 1.0 : 2.0, m == 1 ? 0.0 : -1.0));
      CHK(step(j2rp, j2ci, j2v, t, nullptr, Bm + m * vs, 1.0, 0.0));
    }
    CHK(dot_block_acc<S>(h, A, Bm, (long long)h->n * spc, (long long)(vs / HS), M, first));
    first = false;
  }
  std::vector<double> cm((size_t)M * M * HS);
  HIPCHK(hipMemcpyAsync(cm.data(), h->m2_c.p, cm.size() * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (size_t x = 0; x < (size_t)M * M; x++) {
    mu[2 * x] = cm[x * HS];
    mu[2 * x + 1] = HS == 2 ? cm[x * HS + 1] : 0.0;
  }
  return 0;  // the buffers of moments / autodiff_matrix are untouched
}

template <class TH, class S, bool CX>
int upload_H_t(fkpm_handle* h, DevBuf& dst, const void* val, double mag, double shift) {
  const size_t cnt = (size_t)h->nnzb * h->b * h->b;
  HIPCHK(hipMemcpyAsync(h->hraw.p, val, cnt * sizeof(TH) * (CX ? 2 : 1), hipMemcpyHostToDevice,
                        h->stream));
  return launch(&fkd::k_scale_H<TH, S, CX>, dim3(grid_for(cnt)), dim3(256), h->stream,
                h->hraw.as<const TH>(), dst.as<S>(), h->rowblk.as<const int>(),
                h->ci.as<const int>(), h->nnzb, h->b, mag, shift);
}

template <class S>
int upload_H_s(fkpm_handle* h, DevBuf& dst, const void* val, double mag, double shift) {
  switch (h->dtype) {
    case FKPM_F32: return upload_H_t<float, S, false>(h, dst, val, mag, shift);
    case FKPM_F64: return upload_H_t<double, S, false>(h, dst, val, mag, shift);
    case FKPM_C32: return upload_H_t<float, S, true>(h, dst, val, mag, shift);
    default: return upload_H_t<double, S, true>(h, dst, val, mag, shift);
  }
}

// Pattern upload shared by set_H and energy_scale.
// A periodic 3-D lattice in natural order (block row x + n1 (y + n2 z)),
// recognised from H's block offsets: the three distances |j - i| (mod nb, the
// shorter way round) that at least half the rows have are 1 < n1 < n1 n2 with
// n1 n2 n3 = nb and every side at least 8.  Fills n[3]; false otherwise.
bool lattice_dims(int nb, const int* rp, const int* ci, int n[3]) {
  if (nb < 512) return false;
  std::map<int, long long> cnt;
  for (int i = 0; i < nb; i++)
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      int d = ci[p] - i;
      if (d == 0) continue;
      if (d < 0) d += nb;
      cnt[std::min(d, nb - d)]++;
    }
  std::vector<int> dist;
  for (const auto& kv : cnt)
    if (2 * kv.second >= nb) dist.push_back(kv.first);
  if (dist.size() != 3 || dist[0] != 1) return false;
  const long long n1 = dist[1], c = dist[2];
  if (c % n1 != 0 || nb % c != 0) return false;
  n[0] = (int)n1;
  n[1] = (int)(c / n1);
  n[2] = (int)(nb / c);
  return n[0] >= 8 && n[1] >= 8 && n[2] >= 8;
}

// Sweep order for an n1 x n2 x n3 lattice: one x-slab per XCD (x in
// [floor(k n1 / 8), floor((k+1) n1 / 8)) for XCD k: widths n1/8 rounded down
// or up, about one eighth of the sweep each, the same split into eighths that
// the per-XCD row ranges of the sweep make), each walked in y-tiles of ty rows
// (about 128 sites a plane of the tile) and, inside a tile, plane by plane
// along z.  A row's +-z neighbours are then one tile-plane away in the sweep
// instead of a whole slab-plane (on 64^3: 128 instead of 512 sites), so the
// neighbour rows of both vectors an autodiff step gathers stay in the XCD's
// L2.
bool lattice_order(int nb, const int* rp, const int* ci, std::vector<int>& order) {
  int n[3];
  if (!lattice_dims(nb, rp, ci, n)) return false;
  const int wmax = (n[0] + 7) / 8;
  int ty = 1;
  for (int t = 1; t <= n[1]; t++)
    if (n[1] % t == 0 && wmax * t <= 128) ty = t;
  order.clear();
  order.reserve(nb);
  for (int k = 0; k < 8; k++) {
    const int x0 = (int)((long long)k * n[0] / 8), x1 = (int)((long long)(k + 1) * n[0] / 8);
    for (int y0 = 0; y0 < n[1]; y0 += ty)
      for (int z = 0; z < n[2]; z++)
        for (int yo = 0; yo < ty; yo++)
          for (int x = x0; x < x1; x++) order.push_back(x + n[0] * ((y0 + yo) + n[1] * z));
  }
  return (int)order.size() == nb;
}

int upload_pattern(fkpm_handle* h, int nb, int b, int nnzb, const int* row_ptr,
                   const int* col_idx) {
  h->nb = nb;
  h->b = b;
  h->nnzb = nnzb;
  h->h_rp.assign(row_ptr, row_ptr + nb + 1);
  h->max_deg = 0;
  for (int i = 0; i < nb; i++) h->max_deg = std::max(h->max_deg, row_ptr[i + 1] - row_ptr[i]);
  h->h_ci.assign(col_idx, col_idx + nnzb);
  std::vector<int> rowblk(nnzb > 0 ? nnzb : 1);
  for (int i = 0; i < nb; i++)
    for (int p = row_ptr[i]; p < row_ptr[i + 1]; p++) rowblk[p] = i;
  CHK(h->rp.ensure(sizeof(int) * (nb + 1)));
  CHK(h->ci.ensure(sizeof(int) * (nnzb > 0 ? nnzb : 1)));
  CHK(h->rowblk.ensure(sizeof(int) * rowblk.size()));
  HIPCHK(hipMemcpyAsync(h->rp.p, row_ptr, sizeof(int) * (nb + 1), hipMemcpyHostToDevice, h->stream));
  if (nnzb > 0)
    HIPCHK(hipMemcpyAsync(h->ci.p, col_idx, sizeof(int) * nnzb, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->rowblk.p, rowblk.data(), sizeof(int) * rowblk.size(),
                        hipMemcpyHostToDevice, h->stream));
  const size_t hvb = (size_t)nnzb * b * b * host_elem_bytes(h->dtype);
  CHK(h->hraw.ensure(hvb > 0 ? hvb : 16));
  std::vector<int> order, desc;
  h->has_order = lattice_order(nb, row_ptr, col_idx, order);
  if (h->has_order) {
    desc.resize(3 * (size_t)nb);
    for (int k = 0; k < nb; k++) {
      desc[3 * (size_t)k] = order[k];
      desc[3 * (size_t)k + 1] = row_ptr[order[k]];
      desc[3 * (size_t)k + 2] = row_ptr[order[k] + 1];
    }
    CHK(h->rdesc.ensure(sizeof(int) * desc.size()));
    HIPCHK(hipMemcpyAsync(h->rdesc.p, desc.data(), sizeof(int) * desc.size(),
                          hipMemcpyHostToDevice, h->stream));
  }
  apply_row_order(h);
  // the async copies above read host arrays owned by the caller (and desc):
  // finish them
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

// Symmetric tridiagonal extreme eigenvalues by Sturm-sequence bisection
// (replaces arma::eig_sym on the Lanczos matrix, engine_cusparse.cpp:655-666).
int sturm_below(const std::vector<double>& a, const std::vector<double>& b2, double x) {
  int cnt = 0;
  double q = a[0] - x;
  if (q < 0) cnt++;
  for (size_t j = 1; j < a.size(); j++) {
    if (q == 0) q = 1e-300;
    q = a[j] - x - b2[j] / q;
    if (q < 0) cnt++;
  }
  return cnt;
}

void tridiag_extremes(const std::vector<double>& a, const std::vector<double>& beta,
                      double* emin, double* emax) {
  const size_t n = a.size();
  std::vector<double> b2(n, 0.0);
  double lo = a[0], hi = a[0];
  for (size_t j = 1; j < n; j++) b2[j] = beta[j] * beta[j];
  for (size_t j = 0; j < n; j++) {
    const double r = (j > 0 ? std::fabs(beta[j]) : 0.0) + (j + 1 < n ? std::fabs(beta[j + 1]) : 0.0);
    lo = std::min(lo, a[j] - r);
    hi = std::max(hi, a[j] + r);
  }
  const double pad = 1e-12 * (std::fabs(lo) + std::fabs(hi) + 1.0);
  auto kth = [&](int k) {
    double l = lo - pad, u = hi + pad;
    for (int it = 0; it < 200; it++) {
      const double mid = 0.5 * (l + u);
      if (mid == l || mid == u) break;
      if (sturm_below(a, b2, mid) > k) u = mid; else l = mid;
    }
    return 0.5 * (l + u);
  };
  *emin = kth(0);
  *emax = kth((int)n - 1);
}

// Lanczos on the device in FP64 (engine_cusparse.cpp:598-673).  Uses its own
// buffers, so the engine's H is untouched.  A diagonal block missing from H is
// inserted as an explicit zero block: the fused SpMM kernel takes the own-row
// operand of the dot from the diagonal block.
template <class TH, bool CX>
int lanczos_t(fkpm_handle* h, int nb, int b, int nnzb, const int* row_ptr, const int* col_idx,
              const void* val, double extend, int iters, const void* v_start, double* lo,
              double* hi) {
  fkd::KSet<double> ks;
  const int geoL = 4, geoKQ = 1;
  const bool ok = CX ? fkd::find_kset_dc(b, geoL, geoKQ, &ks) : fkd::find_kset_dr(b, geoL, geoKQ, &ks);
  if (!ok) return fail(FKPM_E_UNSUP, "no Lanczos kernel for block %d", b);
  const int HS = CX ? 2 : 1;
  const int bb = b * b;
  // pattern with every diagonal block present; when H has them all (the
  // usual case) its arrays are uploaded as they are
  int missing = 0;
  for (int i = 0; i < nb; i++) {
    bool diag = false;
    for (int p = row_ptr[i]; p < row_ptr[i + 1] && !diag; p++) diag = (col_idx[p] == i);
    missing += diag ? 0 : 1;
  }
  const bool direct = missing == 0 && std::is_same<TH, double>::value;
  std::vector<int> rp2, ci2;
  std::vector<double> hv2;
  if (!direct) {
    rp2.assign(nb + 1, 0);
    ci2.reserve((size_t)nnzb + missing);
    hv2.reserve(((size_t)nnzb + missing) * bb * HS);
    for (int i = 0; i < nb; i++) {
      bool diag = false;
      for (int p = row_ptr[i]; p < row_ptr[i + 1]; p++) diag |= (col_idx[p] == i);
      for (int p = row_ptr[i]; p < row_ptr[i + 1]; p++) {
        ci2.push_back(col_idx[p]);
        const TH* v = static_cast<const TH*>(val) + (size_t)p * bb * HS;
        for (int e = 0; e < bb * HS; e++) hv2.push_back((double)v[e]);
      }
      if (!diag) {
        ci2.push_back(i);
        for (int e = 0; e < bb * HS; e++) hv2.push_back(0.0);
      }
      rp2[i + 1] = (int)ci2.size();
    }
  }
  const int* up_rp = direct ? row_ptr : rp2.data();
  const int* up_ci = direct ? col_idx : ci2.data();
  const double* up_hv = direct ? static_cast<const double*>(val) : hv2.data();
  const int nnz2 = direct ? nnzb : (int)ci2.size();
  const int n = nb * b;
  const int qp = geoL * geoKQ;  // one probe column padded to one lane-group pass
  const int sp = qp * (2 / HS);
  const size_t len = (size_t)n * sp * HS;  // doubles per vector
  const int gwg = (fkd::WG / fkd::WAVE) * (fkd::WAVE / geoL);
  int nwg = std::min((nb + gwg - 1) / gwg, 1024);
  int chunk = (nb + nwg - 1) / nwg;
  chunk = ((chunk + gwg - 1) / gwg) * gwg;
  nwg = (nb + chunk - 1) / chunk;

  struct Bufs {
    DevBuf rp, ci, hv, v0, v1, w, part, ps, coef, used;
    ~Bufs() {
      DevBuf* all[] = {&rp, &ci, &hv, &v0, &v1, &w, &part, &ps, &coef, &used};
      for (DevBuf* d : all) d->release();
    }
  } B;
  const size_t hv_count = (size_t)nnz2 * bb * HS;
  CHK(B.rp.ensure(sizeof(int) * (nb + 1)));
  CHK(B.ci.ensure(sizeof(int) * std::max(nnz2, 1)));
  CHK(B.hv.ensure(sizeof(double) * std::max<size_t>(hv_count, 1)));
  CHK(B.v0.ensure(len * sizeof(double)));
  CHK(B.v1.ensure(len * sizeof(double)));
  CHK(B.w.ensure(len * sizeof(double)));
  CHK(B.part.ensure((size_t)nwg * 2 * sizeof(double)));
  CHK(B.ps.ensure(2 * sizeof(double)));
  HIPCHK(hipMemcpyAsync(B.rp.p, up_rp, sizeof(int) * (nb + 1), hipMemcpyHostToDevice, h->stream));
  if (nnz2 > 0) {
    HIPCHK(hipMemcpyAsync(B.ci.p, up_ci, sizeof(int) * nnz2, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(B.hv.p, up_hv, sizeof(double) * hv_count, hipMemcpyHostToDevice,
                          h->stream));
  }
  CHK(zero_doubles(h, B.v0.p, len));
  CHK(zero_doubles(h, B.w.p, len));
  {
    std::vector<double> packed(len, 0.0);
    double nrm = 0.0;
    for (int r = 0; r < n; r++)
      for (int c = 0; c < HS; c++) {
        const double x = (double)static_cast<const TH*>(v_start)[(size_t)r * HS + c];
        nrm += x * x;
      }
    nrm = std::sqrt(nrm);
    if (!(nrm > 0) || !std::isfinite(nrm)) return fail(FKPM_E_ARG, "energy_scale: bad start vector");
    for (int r = 0; r < n; r++)
      for (int c = 0; c < HS; c++)
        packed[(size_t)r * sp * HS + c] = (double)static_cast<const TH*>(v_start)[(size_t)r * HS + c] / nrm;
    HIPCHK(hipMemcpyAsync(B.v1.p, packed.data(), len * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));  // host staging vectors go out of scope
  }
  // the iteration runs on the stream without a host round trip: alpha, beta
  // and the invariant-subspace cut stay on the device until the end
  const dim3 grid(nwg), block(fkd::WG);
  CHK(B.coef.ensure(2 * (size_t)iters * sizeof(double)));
  CHK(B.used.ensure(sizeof(int)));
  double* coef = static_cast<double*>(B.coef.p);
  int* dused = static_cast<int*>(B.used.p);
  const double* dps = static_cast<const double*>(B.ps.p);
  auto spmv_dot = [&]() -> int {
    // w = H v1; ps = (|v1|^2, Re <w, v1>)
    CHK(launch(ks.mom, grid, block, h->stream, (const int*)B.rp.p, (const int*)B.ci.p,
               (const double*)B.hv.p, (const double*)B.v1.p, (const double*)nullptr,
               (double*)B.w.p, nb, qp, chunk, 1.0, 0.0, (double*)B.part.p, 0, no_rsource<double>()));
    return launch(&fkd::k_reduce_partials, dim3(1), block, h->stream, (const double*)B.part.p,
                  nwg, (double*)B.ps.p);
  };
  CHK(launch(&fkd::k_lanczos_edge, dim3(1), dim3(64), h->stream, dps, coef, iters, dused, 0));
  for (int j = 1; j < iters; j++) {
    CHK(spmv_dot());
    CHK(launch(&fkd::k_lanczos_orth, dim3(nwg), block, h->stream, (double*)B.w.p,
               (const double*)B.v1.p, (const double*)B.v0.p, len, dps, coef, iters, j,
               (double*)B.part.p));
    CHK(launch(&fkd::k_reduce_partials, dim3(1), block, h->stream, (const double*)B.part.p, nwg,
               (double*)B.ps.p));
    CHK(launch(&fkd::k_lanczos_next, dim3(grid_for(len)), dim3(256), h->stream, (double*)B.v0.p,
               (double*)B.v1.p, (const double*)B.w.p, len, dps, coef, iters, j, dused));
  }
  CHK(spmv_dot());
  CHK(launch(&fkd::k_lanczos_edge, dim3(1), dim3(64), h->stream, dps, coef, iters, dused, 1));
  std::vector<double> cf(2 * (size_t)iters);
  int used = iters;
  HIPCHK(hipMemcpyAsync(cf.data(), coef, cf.size() * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipMemcpyAsync(&used, dused, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  std::vector<double> alpha(cf.begin(), cf.begin() + used), beta(cf.begin() + iters,
                                                                 cf.begin() + iters + used);
  double emin = 0, emax = 0;
  tridiag_extremes(alpha, beta, &emin, &emax);
  const double slack = extend * (emax - emin);
  *lo = emin - slack;
  *hi = emax + slack;
  return 0;
}


}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* fkpm_last_error(void) { return g_last_error.c_str(); }

int fkpm_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

int fkpm_dev_open(int device, int dtype, int prec, fkpm_handle** out) {
  if (!out) return fail(FKPM_E_ARG, "null handle pointer");
  *out = nullptr;
  if (dtype < FKPM_F32 || dtype > FKPM_C64) return fail(FKPM_E_ARG, "bad dtype %d", dtype);
  if (prec != FKPM_PREC_FP64 && prec != FKPM_PREC_FP32) return fail(FKPM_E_ARG, "bad prec %d", prec);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(FKPM_E_NODEV, "no HIP device is available");
  if (device < 0 || device >= count)
    return fail(FKPM_E_NODEV, "invalid device #%d (%d available)", device, count);
  HIPCHK(hipSetDevice(device));
  fkpm_handle* h = new fkpm_handle();
  h->device = device;
  h->dtype = dtype;
  h->prec = prec;
  h->cx = dtype_complex(dtype);
  h->single = (prec == FKPM_PREC_FP32);
  // FKPM_AUTOTUNE=1: tune at the first use of a shape (FKPM_OPT_AUTOTUNE for
  // callers that cannot reach fkpm_set_option, such as the C++ API)
  if (const char* e = getenv("FKPM_AUTOTUNE")) h->autotune_on_use = e[0] != '\0' && e[0] != '0';
  // FKPM_GRAPH=0/1/-1: the FKPM_OPT_GRAPH default for handles opened afterwards
  if (const char* e = getenv("FKPM_GRAPH")) {
    const int v = atoi(e);
    if (v >= -1 && v <= 1) h->graph_mode = v;
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
    delete h;
    return fail(FKPM_E_HIP, "stream/event creation failed");
  }
  *out = h;
  return 0;
}

int fkpm_dev_close(fkpm_handle* h) {
  if (!h) return 0;
  delete h;
  return 0;
}

void* fkpm_stream(fkpm_handle* h) { return h ? (void*)h->stream : nullptr; }

int fkpm_synchronize(fkpm_handle* h) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int fkpm_set_H(fkpm_handle* h, int nb, int b, int nnzb, const int* row_ptr, const int* col_idx,
               const void* val, double avg, double mag) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  HIPCHK(hipSetDevice(h->device));
  if (b <= 0) return fail(FKPM_E_ARG, "bad block size %d", b);
  if (!fkd::block_supported(b)) return fail(FKPM_E_UNSUP, "block size %d has no kernels", b);
  if (!(mag > 0) || !std::isfinite(avg)) return fail(FKPM_E_ARG, "bad energy scale avg=%g mag=%g", avg, mag);
  if (nnzb > 0 && !val) return fail(FKPM_E_ARG, "null H values");
  CHK(check_pattern("set_H", nb, nnzb, row_ptr, nullptr, col_idx));
  for (int i = 0; i < nb; i++) {  // engine_cusparse.cpp:716-727: every diagonal block present
    bool found = false;
    for (int p = row_ptr[i]; p < row_ptr[i + 1]; p++) found |= (col_idx[p] == i);
    if (!found) return fail(FKPM_E_ARG, "set_H: diagonal block (%d,%d) is missing", i, i);
  }
  // The kernel choice depends on the block size, the block-row and block
  // counts (grids, tuned-shape key) and on whether every row has at most
  // kRowResidentMax blocks (row-resident matrix-core kernels read only that many),
  // or at most 7 (their instantiation without a padding slot): a change of any
  // of them picks the kernels again at the next use.
  const int old_b = h->b, old_nb = h->nb, old_nnzb = h->nnzb, old_deg = h->max_deg;
  CHK(upload_pattern(h, nb, b, nnzb, row_ptr, col_idx));
  const bool geo_change = b != old_b || nb != old_nb || nnzb != old_nnzb ||
                          (old_deg <= kRowResidentMax) != (h->max_deg <= kRowResidentMax) ||
                          (old_deg <= 7) != (h->max_deg <= 7);
  h->avg = avg;
  h->mag = mag;
  const size_t eb = h->single ? 4 : 8;
  CHK(h->hv.ensure((size_t)(nnzb > 0 ? nnzb : 1) * b * b * (h->cx ? 2 : 1) * eb));
  if (h->single)
    CHK(upload_H_s<float>(h, h->hv, val, mag, avg / mag));
  else
    CHK(upload_H_s<double>(h, h->hv, val, mag, avg / mag));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->h_set = true;
  h->alpha_M = 0;
  h->tab_K = 0;
  if (geo_change) h->r_dirty = true;
  return 0;
}

int fkpm_set_R(fkpm_handle* h, int n, int s, const void* R_colmajor) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  HIPCHK(hipSetDevice(h->device));
  if (n <= 0 || s <= 0 || !R_colmajor)  // engine_cusparse.cpp:678-681
    return fail(FKPM_E_ARG, "cannot handle an empty R matrix (n=%d, s=%d)", n, s);
  const size_t bytes = (size_t)n * s * host_elem_bytes(h->dtype);
  CHK(h->rraw.ensure(bytes));
  HIPCHK(hipMemcpyAsync(h->rraw.p, R_colmajor, bytes, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->n = n;
  h->s = s;
  h->r_set = true;
  h->r_dirty = true;
  h->alpha_M = 0;
  h->tab_K = 0;
  return 0;
}

int fkpm_moments(fkpm_handle* h, int M, double* mu) {
  if (!h || !mu) return fail(FKPM_E_ARG, "null argument");
  HIPCHK(hipSetDevice(h->device));
  if (M < 2 || M % 2 != 0) return fail(FKPM_E_ARG, "moments: M=%d must be even and >= 2", M);
  CHK(ensure_ready(h));
  h->alpha_M = 0;
  if (h->single) return moments_t<float>(h, M, mu);
  return moments_t<double>(h, M, mu);
}

int fkpm_set_D_pattern(fkpm_handle* h, int nbD, int nnzbD, const int* row_ptr, const int* row_idx,
                       const int* col_idx) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  HIPCHK(hipSetDevice(h->device));
  CHK(check_pattern("D pattern", nbD, nnzbD, row_ptr, row_idx, col_idx));
  // block (i,j) -> index of block (j,i), or -1
  std::vector<int> tpos(nnzbD > 0 ? nnzbD : 1, -1);
  bool sym = true;
  for (int i = 0; i < nbD; i++)
    for (int p = row_ptr[i]; p < row_ptr[i + 1]; p++) {
      const int j = col_idx[p];
      int t = -1;
      for (int q = row_ptr[j]; q < row_ptr[j + 1]; q++)
        if (col_idx[q] == i) {
          t = q;
          break;
        }
      tpos[p] = t;
      if (t < 0 && i >= j) sym = false;  // the reference aborts only for these
    }
  CHK(h->drp.ensure(sizeof(int) * (nbD + 1)));
  CHK(h->dci.ensure(sizeof(int) * (nnzbD > 0 ? nnzbD : 1)));
  CHK(h->tpos.ensure(sizeof(int) * tpos.size()));
  HIPCHK(hipMemcpyAsync(h->drp.p, row_ptr, sizeof(int) * (nbD + 1), hipMemcpyHostToDevice, h->stream));
  if (nnzbD > 0)
    HIPCHK(hipMemcpyAsync(h->dci.p, col_idx, sizeof(int) * nnzbD, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->tpos.p, tpos.data(), sizeof(int) * tpos.size(), hipMemcpyHostToDevice,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->nbD = nbD;
  h->nnzbD = nnzbD;
  h->hd_rp.assign(row_ptr, row_ptr + nbD + 1);
  h->hd_ci.assign(col_idx, col_idx + nnzbD);
  h->d_set = true;
  h->d_sym = sym;
  h->d_valid = false;
  return 0;
}

int fkpm_stoch_matrix(fkpm_handle* h, int M, const double* c, void* D_val) {
  if (!h || !c) return fail(FKPM_E_ARG, "null argument");
  HIPCHK(hipSetDevice(h->device));
  if (M < 2) return fail(FKPM_E_ARG, "stoch_matrix: need at least 2 coefficients (M=%d)", M);
  CHK(ensure_ready(h));
  CHK(check_D_ready(h));
  h->d_valid = false;
  if (h->single)
    CHK(stoch_t<float>(h, M, c));
  else
    CHK(stoch_t<double>(h, M, c));
  h->d_valid = true;
  if (D_val) return copy_D_out(h, D_val);
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int fkpm_autodiff_matrix(fkpm_handle* h, int M, const double* c, void* D_val) {
  if (!h || !c) return fail(FKPM_E_ARG, "null argument");
  HIPCHK(hipSetDevice(h->device));
  if (M < 2 || M % 2 != 0) return fail(FKPM_E_ARG, "autodiff_matrix: M=%d must be even and >= 2", M);
  if (h->alpha_M != M)
    return fail(FKPM_E_STATE,
                "autodiff_matrix requires moments(M) with the same M (%d) immediately before "
                "(held: %d)", M, h->alpha_M);
  CHK(check_D_ready(h));
  if (!h->d_sym)
    return fail(FKPM_E_ARG, "autodiff_matrix: a block (i,j) with i >= j of D has no (j,i) partner");
  h->d_valid = false;
  if (h->single)
    CHK(autodiff_t<float>(h, M, c));
  else
    CHK(autodiff_t<double>(h, M, c));
  h->d_valid = true;
  if (D_val) return copy_D_out(h, D_val);
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int fkpm_moments2(fkpm_handle* h, int M, int nnzb1, const int* rp1, const int* ci1,
                  const void* v1, int nnzb2, const int* rp2, const int* ci2, const void* v2,
                  int col_chunk, double* mu) {
  if (!h || !mu || !rp1 || !ci1 || !v1 || !rp2 || !ci2 || !v2)
    return fail(FKPM_E_ARG, "moments2: null argument");
  if (M < 2 || M % 2 != 0) return fail(FKPM_E_ARG, "moments2: M = %d must be even and >= 2", M);
  HIPCHK(hipSetDevice(h->device));
  CHK(ensure_ready(h));
  CHK(check_pattern("j1", h->nb, nnzb1, rp1, nullptr, ci1));
  CHK(check_pattern("j2", h->nb, nnzb2, rp2, nullptr, ci2));
  if (h->single) {
    CHK(upload_current_s<float>(h, 0, nnzb1, rp1, ci1, v1));
    CHK(upload_current_s<float>(h, 1, nnzb2, rp2, ci2, v2));
  } else {
    CHK(upload_current_s<double>(h, 0, nnzb1, rp1, ci1, v1));
    CHK(upload_current_s<double>(h, 1, nnzb2, rp2, ci2, v2));
  }
  // vector budget: half the free device memory, at most 64 GiB, unless the
  // environment names one
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b));
  size_t budget = std::min((size_t)64 << 30, free_b / 2);
  if (const char* e = getenv("FKPM_MOMENTS2_BYTES")) budget = (size_t)strtoull(e, nullptr, 10);
  return h->single ? moments2_t<float>(h, M, col_chunk, budget, mu)
                   : moments2_t<double>(h, M, col_chunk, budget, mu);
}

int fkpm_get_D(fkpm_handle* h, void* D_val) {
  if (!h || !D_val) return fail(FKPM_E_ARG, "null argument");
  HIPCHK(hipSetDevice(h->device));
  if (!h->d_valid) return fail(FKPM_E_STATE, "no D has been computed");
  return copy_D_out(h, D_val);
}

int fkpm_energy_scale(fkpm_handle* h, int nb, int b, int nnzb, const int* row_ptr,
                      const int* col_idx, const void* val, double extend, int iters,
                      const void* v_start, double* lo, double* hi) {
  if (!h || !v_start || !lo || !hi) return fail(FKPM_E_ARG, "null argument");
  HIPCHK(hipSetDevice(h->device));
  if (b <= 0 || !fkd::block_supported(b)) return fail(FKPM_E_UNSUP, "block size %d has no kernels", b);
  if (iters < 1) return fail(FKPM_E_ARG, "energy_scale: iters=%d", iters);
  if (nnzb > 0 && !val) return fail(FKPM_E_ARG, "null H values");
  CHK(check_pattern("energy_scale", nb, nnzb, row_ptr, nullptr, col_idx));
  switch (h->dtype) {
    case FKPM_F32: return lanczos_t<float, false>(h, nb, b, nnzb, row_ptr, col_idx, val, extend, iters, v_start, lo, hi);
    case FKPM_F64: return lanczos_t<double, false>(h, nb, b, nnzb, row_ptr, col_idx, val, extend, iters, v_start, lo, hi);
    case FKPM_C32: return lanczos_t<float, true>(h, nb, b, nnzb, row_ptr, col_idx, val, extend, iters, v_start, lo, hi);
    default: return lanczos_t<double, true>(h, nb, b, nnzb, row_ptr, col_idx, val, extend, iters, v_start, lo, hi);
  }
}

int fkpm_set_timing(fkpm_handle* h, int on) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  h->timing = on != 0;
  return 0;
}

int fkpm_get_timing(fkpm_handle* h, double* ms_moments, int* n_moments, double* ms_stoch,
                    int* n_stoch, double* ms_autodiff, int* n_autodiff) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  if (ms_moments) *ms_moments = h->ms_mom;
  if (n_moments) *n_moments = h->n_mom;
  if (ms_stoch) *ms_stoch = h->ms_stoch;
  if (n_stoch) *n_stoch = h->n_stoch;
  if (ms_autodiff) *ms_autodiff = h->ms_ad;
  if (n_autodiff) *n_autodiff = h->n_ad;
  return 0;
}

int fkpm_get_config(fkpm_handle* h, int* lanes, int* kq, int* s_padded, int* n_wg,
                    int* fused_autodiff) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  if (lanes) *lanes = h->kc[K_MOM].L;
  if (kq) *kq = h->kc[K_MOM].KQ;
  if (s_padded) *s_padded = h->sp;
  if (n_wg) *n_wg = h->kc[K_MOM].nwg;
  if (fused_autodiff) *fused_autodiff = h->d_fused ? 1 : 0;
  return 0;
}

int fkpm_get_kernel_config(fkpm_handle* h, int kind, int* lanes, int* kq, int* n_wg, int* nt) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  if (kind < 0 || kind >= NKIND) return fail(FKPM_E_ARG, "bad kernel kind %d", kind);
  const KCfg& c = h->kc[kind];
  if (lanes) *lanes = c.L;
  if (kq) *kq = c.KQ;
  if (n_wg) *n_wg = c.nwg;
  if (nt) *nt = c.nt;
  return 0;
}

int fkpm_get_kernel_mfma(fkpm_handle* h, int kind, int* mfma) {
  if (!h || !mfma) return fail(FKPM_E_ARG, "null argument");
  if (kind < 0 || kind >= NKIND) return fail(FKPM_E_ARG, "bad kernel kind %d", kind);
  *mfma = h->kc[kind].mfma;
  return 0;
}

int fkpm_get_alpha_table(fkpm_handle* h, int* stored) {
  if (!h || !stored) return fail(FKPM_E_ARG, "null argument");
  *stored = (h->a_stored && h->alpha_M > 0) ? 1 : 0;
  return 0;
}

int fkpm_get_stoch_method(fkpm_handle* h, int* method) {
  if (!h || !method) return fail(FKPM_E_ARG, "null argument");
  *method = h->stoch_used;
  return 0;
}

int fkpm_get_kernel_sweep(fkpm_handle* h, int kind, int* sweep) {
  if (!h || !sweep) return fail(FKPM_E_ARG, "null argument");
  if (kind < 0 || kind >= NKIND) return fail(FKPM_E_ARG, "bad kernel kind %d", kind);
  *sweep = h->kc[kind].sweep;
  return 0;
}

int fkpm_lattice_order(int nb, int nnzb, const int* row_ptr, const int* col_idx, int* order,
                       int* dims) {
  if (nb <= 0 || nnzb < 0 || !row_ptr || (nnzb > 0 && !col_idx) || !order || !dims)
    return fail(FKPM_E_ARG, "lattice_order: bad arguments");
  CHK(check_pattern("lattice_order", nb, nnzb, row_ptr, nullptr, col_idx));
  std::vector<int> ord;
  if (!lattice_order(nb, row_ptr, col_idx, ord)) return 0;
  int n[3];
  lattice_dims(nb, row_ptr, col_idx, n);
  std::copy(ord.begin(), ord.end(), order);
  std::copy(n, n + 3, dims);
  return 1;
}

int fkpm_get_graph_stats(fkpm_handle* h, long long* launches, long long* captures) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  if (launches) *launches = h->n_graph_launches;
  if (captures) *captures = h->n_graph_captures;
  return 0;
}

int fkpm_get_state(fkpm_handle* h, int* r_compact, int* tuned, int* moments_M) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  if (r_compact) *r_compact = h->r_compact ? 1 : 0;
  if (tuned) *tuned = h->tuned ? 1 : 0;
  if (moments_M) *moments_M = h->alpha_M;
  return 0;
}

int fkpm_set_geometry(fkpm_handle* h, int lanes, int kq, int max_wg) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  if (lanes < 0 || kq < 0 || max_wg < 0) return fail(FKPM_E_ARG, "bad geometry");
  if (lanes != 0 && h->b > 0 && !compiled(h->b, lanes, kq > 0 ? kq : 1))
    return fail(FKPM_E_UNSUP, "geometry lanes=%d kq=%d not compiled for block %d", lanes, kq, h->b);
  h->force_L = lanes;
  h->force_KQ = kq;
  h->max_wg = max_wg;
  h->r_dirty = true;
  h->alpha_M = 0;
  return 0;
}

int fkpm_autotune(fkpm_handle* h) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  HIPCHK(hipSetDevice(h->device));
  CHK(ensure_ready(h));
  if (h->force_L > 0) return fail(FKPM_E_STATE, "a forced geometry is set (fkpm_set_geometry)");
  CHK(tune(h));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int fkpm_set_option(fkpm_handle* h, int option, int value) {
  if (!h) return fail(FKPM_E_ARG, "null handle");
  switch (option) {
    case FKPM_OPT_STOCH_METHOD:
      if (value != FKPM_STOCH_CLENSHAW && value != FKPM_STOCH_FORWARD && value != FKPM_STOCH_AUTO)
        return fail(FKPM_E_ARG, "bad stoch_matrix method %d", value);
      h->stoch_method = value;
      return 0;
    case FKPM_OPT_AUTOTUNE:
      h->autotune_on_use = value != 0;
      return 0;
    case FKPM_OPT_STORE_ALPHA:
      if (value < 0 || value > 2) return fail(FKPM_E_ARG, "bad store-alpha mode %d", value);
      h->store_alpha = value;
      h->alpha_M = 0;  // moments(M) again before autodiff_matrix
      h->a_stored = false;
      return 0;
    case FKPM_OPT_SWEEP:
      if (value < -1 || value > 64) return fail(FKPM_E_ARG, "bad sweep %d", value);
      h->sweep_force = value;
      if (value >= 0)
        for (int k = 0; k < NKIND; k++)
          if (h->kc[k].L > 0) {
            h->kc[k].sweep = (h->kc[k].nt & fkd::NT_ORDER) ? 0 : value;
            derive_grid(h->nb, h->kc[k]);
          }
      return 0;
    case FKPM_OPT_GRAPH:
      if (value < -1 || value > 1) return fail(FKPM_E_ARG, "bad graph switch %d", value);
      h->graph_mode = value;
      return 0;
    case FKPM_OPT_SLICE:
      if (value < -1 || value > 1) return fail(FKPM_E_ARG, "bad slice switch %d", value);
      h->slice_force = value;
      if (value >= 0)
        for (int k = 0; k < NKIND; k++)
          if (h->kc[k].L > 0 && !h->kc[k].mfma)
            h->kc[k].nt = (h->kc[k].nt & ~fkd::NT_SLICE) | (value ? fkd::NT_SLICE : 0);
      return 0;
    case FKPM_OPT_ROW_ORDER:
      if (value < -1 || value > 1) return fail(FKPM_E_ARG, "bad row-order switch %d", value);
      h->order_force = value;
      apply_row_order(h);
      return 0;
    case FKPM_OPT_ALPHA_TABLE_MB:
      if (value < -1) return fail(FKPM_E_ARG, "bad table budget %d MiB", value);
      h->table_cap_mb = value;
      return 0;
    case FKPM_OPT_MFMA:
      if (value < 0 || value > 2) return fail(FKPM_E_ARG, "bad MFMA switch %d", value);
      if (h->mfma_on != (value != 0) || h->rr_force != (value == 2)) {
        h->mfma_on = value != 0;
        h->rr_force = value == 2;
        h->r_dirty = true;  // geometry (and padded width) chosen again at next use
        h->alpha_M = 0;
        h->tuned = false;
      }
      return 0;
  }
  return fail(FKPM_E_ARG, "unknown option %d", option);
}

int fkpm_bandwidth_probe(fkpm_handle* h, size_t bytes, int reps, double* gbs) {
  if (!h || !gbs || bytes == 0 || bytes % 16 != 0 || reps < 1)
    return fail(FKPM_E_ARG, "bandwidth_probe: bad arguments");
  HIPCHK(hipSetDevice(h->device));
  DevBuf a, b;
  struct Rel {
    DevBuf *x, *y;
    ~Rel() {
      x->release();
      y->release();
    }
  } rel{&a, &b};
  CHK(a.ensure(bytes));
  CHK(b.ensure(bytes));
  const size_t n = bytes / 16;
  CHK(zero_doubles(h, a.p, n * 2));
  // the plain copy (1024 workgroups), then the streaming one (4096); the
  // faster of the two is the ceiling
  double best = 0.0;
  for (int v = 0; v < 2; v++) {
    auto fn = v == 0 ? &fkd::k_copy16 : &fkd::k_copy16_nt;
    const int grid = v == 0 ? 1024 : 4096;
    CHK(launch(fn, dim3(grid), dim3(fkd::WG), h->stream, (const double2*)a.p, (double2*)b.p, n));
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int r = 0; r < reps; r++)
      CHK(launch(fn, dim3(grid), dim3(fkd::WG), h->stream, (const double2*)a.p, (double2*)b.p,
                 n));
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    best = std::max(best, 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9);
  }
  *gbs = best;
  return 0;
}

}  // extern "C"

namespace fkd {
int geometries_for_block(int b, GeomChoice* out, int cap) {
  int n = 0;
#define FKD_GEO(bb, l, kq) \
  if (bb == b && n < cap) out[n++] = GeomChoice{l, kq};
  FKD_FOR_EACH_GEOMETRY(FKD_GEO)
#undef FKD_GEO
  return n;
}
}  // namespace fkd

// ---------------------------------------------------------------------------
// Host functions for the batched engine (kpm_batch.hip).
// ---------------------------------------------------------------------------
namespace fkpm_internal {

int scale_H_cx(hipStream_t st, const double* raw, void* dst, bool single, const int* row_of_block,
               const int* ci, int nnzb, int b, double mag, double shift) {
  const size_t cnt = (size_t)nnzb * b * b;
  if (cnt == 0) return 0;
  if (single)
    return launch(&fkd::k_scale_H<double, float, true>, dim3(grid_for(cnt)), dim3(256), st, raw,
                  (float*)dst, row_of_block, ci, nnzb, b, mag, shift);
  return launch(&fkd::k_scale_H<double, double, true>, dim3(grid_for(cnt)), dim3(256), st, raw,
                (double*)dst, row_of_block, ci, nnzb, b, mag, shift);
}

int pack_R_cx(hipStream_t st, const double* raw, void* dst, bool single, int n, int s, int sp) {
  const dim3 grid((n + 63) / 64, (sp + 15) / 16);
  if (single)
    return launch(&fkd::k_pack_R<double, float, true>, grid, dim3(fkd::WG), st, raw, (float*)dst, n,
                  s, sp);
  return launch(&fkd::k_pack_R<double, double, true>, grid, dim3(fkd::WG), st, raw, (double*)dst,
                n, s, sp);
}

int set_error(int code, const char* msg) { return fail(code, "%s", msg); }

bool lattice_row_order(int nb, const int* rp, const int* ci, std::vector<int>& order) {
  return lattice_order(nb, rp, ci, order);
}

}  // namespace fkpm_internal
