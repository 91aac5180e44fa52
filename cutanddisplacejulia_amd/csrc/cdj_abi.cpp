// C ABI of the engine (include/cdj.h): contexts, device buffers, host<->device staging and dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cdj.h"
#include "batch.hpp"
#include "crack_tip.hpp"
#include "friction.hpp"
#include "learned.hpp"
#include "linalg.hpp"
#include "mixed.hpp"
#include "td_kernels.hpp"
#include "univariate.hpp"

using namespace cdj;

struct cdj_mlp {
  cdj::Mlp m;
};

struct BatchPlan;
void batch_plan_free(BatchPlan* p);

// Scratch of cdj_assemble_dev_noalloc: a fixed device array, one copy per GPU, held until the process ends.  In host
// code `scratch` is only a stand-in for the symbol; its device address on the current GPU comes from
// hipGetSymbolAddress (noalloc_scratch_ptr) and only that address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[CDJ_NOALLOC_SCRATCH_BYTES];

struct cdj_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream of the blocked LU's look-ahead (non-blocking, default priority; used only with `stream`-ordered
  // work through events)
  hipStream_t side = nullptr;
  std::string err;
  // grow-only device scratch for per-call set-up records (triangle and receiver records)
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // grow-only device scratch of the half-space assembly: edge topology and edge records, edge responses
  void* scratch_topo = nullptr;
  size_t scratch_topo_bytes = 0;
  void* scratch_resp = nullptr;
  size_t scratch_resp_bytes = 0;
  // HIP events around the last assembly / TD kernel launch, on the stream it ran on; ev2..ev3 around the edge
  // responses of the last half-space assembly (when it ran them in one launch)
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  bool timed = false;
  bool edge_timed = false;
  int64_t last_edges = 0;
  // entries recomputed in FP64 by the last mixed assembly
  int64_t near_entries = 0;
  // device buffers, tuned tiles and captured graph of the last batch shape (cdj_batched_slip)
  BatchPlan* batch = nullptr;
  // this device's copy of `scratch` (cdj_assemble_dev_noalloc), shared by every context on the device
  char* noalloc_scratch = nullptr;
};

namespace {

thread_local std::string g_create_err;

// Device address of `scratch` on `device` (which must be current), looked up once per device.
hipError_t noalloc_scratch_ptr(int device, char** out) {
  static std::mutex mu;
  static std::map<int, char*> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it == addr.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    it = addr.emplace(device, static_cast<char*>(p)).first;
  }
  *out = it->second;
  return hipSuccess;
}

int fail(cdj_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(cdj_ctx* c, hipError_t e, const char* where) {
  return fail(c, e == hipErrorOutOfMemory ? CDJ_ERR_NOMEM : CDJ_ERR_DEVICE,
              std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(expr)                                   \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_fail(ctx, _e, #expr); \
  } while (0)

// Scoped device allocation.
struct DBuf {
  void* p = nullptr;
  ~DBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  hipError_t alloc(size_t bytes) {
    if (bytes == 0) bytes = 8;
    return hipMalloc(&p, bytes);
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Grow-only device buffer (`buf`, `have` bytes) of at least `bytes`; the previous contents are not kept.
hipError_t grow(void** buf, size_t* have, size_t bytes) {
  if (bytes > *have) {
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    hipError_t e = hipMalloc(buf, bytes);
    if (e != hipSuccess) return e;
    *have = bytes;
  }
  return hipSuccess;
}

// Device scratch of at least `bytes` owned by the context (not valid across overlapping calls).
hipError_t ctx_scratch(cdj_ctx* c, size_t bytes, void** out) {
  hipError_t e = grow(&c->scratch, &c->scratch_bytes, bytes);
  *out = c->scratch;
  return e;
}

// Edge responses are kept for at most this many bytes at once; larger problems take the receivers in chunks.
constexpr size_t kEdgeResponseBytes = size_t(4) << 30;

// FP64 half-space traction entries of nrecv receivers (device records `recv`) against the n source triangles
// (device corners P1..P3, records `rec` built), written through `out`.  The free-surface part is evaluated once per
// mesh edge and shared by the edge's triangles (tdk_edge_response, tdk_traction_combine); when a triangle lies in
// z = 0 the fused per-pair kernel runs instead, since its image pass needs the accumulated strain.  With `timed`,
// ev0..ev1 bracket the assembly kernels (and ev2..ev3 the edge responses when they run in one launch).
int traction_hs_fp64(cdj_ctx* ctx, const double* recv, int64_t nrecv, const double* P1, const double* P2,
                     const double* P3, const double* rec, int64_t n, int tpb, double nu, double lam, double G,
                     const OutPtrs<double>& out, bool timed, hipStream_t st) {
  ctx->edge_timed = false;
  ctx->last_edges = 0;
  if (nrecv <= 0 || n <= 0) return CDJ_OK;
  const uint64_t S = edge_table_slots(n);
  const size_t ints = 2 * S + 4 * 3 * (size_t)n + 2;
  const size_t topo_bytes = ((sizeof(int) * ints + 255) / 256) * 256;
  HIPCHK(grow(&ctx->scratch_topo, &ctx->scratch_topo_bytes, topo_bytes + sizeof(double) * REC_EDGE * 3 * n));
  int* ip = static_cast<int*>(ctx->scratch_topo);
  EdgeTopo topo{ip, ip + S, ip + 2 * S, ip + 2 * S + 3 * n, ip + 2 * S + 6 * n, ip + 2 * S + 9 * n,
                ip + 2 * S + 12 * n};
  double* erec = reinterpret_cast<double*>(static_cast<char*>(ctx->scratch_topo) + topo_bytes);
  launch_edge_topology(P1, P2, P3, n, topo, st);
  int cnt[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(cnt, topo.count, sizeof(cnt), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const int64_t E = cnt[0];
  if (cnt[1] || E <= 0) {
    if (timed) HIPCHK(hipEventRecord(ctx->ev0, st));
    launch_traction<double>(recv, nrecv, rec, n, 1, tpb, nu, lam, G, out, st);
    if (timed) HIPCHK(hipEventRecord(ctx->ev1, st));
    HIPCHK(hipGetLastError());
    return CDJ_OK;
  }
  ctx->last_edges = E;
  int64_t chunk = (int64_t)(kEdgeResponseBytes / (sizeof(double) * 9 * (size_t)E));
  chunk = std::max<int64_t>(256, chunk / 256 * 256);
  chunk = std::min<int64_t>(chunk, nrecv);
  HIPCHK(grow(&ctx->scratch_resp, &ctx->scratch_resp_bytes, sizeof(double) * 9 * (size_t)E * chunk));
  double* Tb = static_cast<double*>(ctx->scratch_resp);
  if (timed) HIPCHK(hipEventRecord(ctx->ev0, st));
  launch_edge_setup(P1, P2, P3, n, rec, topo, erec, st);
  const bool one = chunk == nrecv;
  for (int64_t r0 = 0; r0 < nrecv; r0 += chunk) {
    const int64_t m = std::min<int64_t>(chunk, nrecv - r0);
    if (timed && one) HIPCHK(hipEventRecord(ctx->ev2, st));
    launch_edge_response(recv + r0 * RECV, m, erec, E, nu, lam, G, Tb, m, st);
    if (timed && one) HIPCHK(hipEventRecord(ctx->ev3, st));
    OutPtrs<double> o = out;
    for (int k = 0; k < 9; ++k) o.p[k] = out.p[k] + r0;
    launch_traction_combine(recv + r0 * RECV, m, rec, n, tpb, nu, lam, G, topo.side_edge, Tb, m, o, st);
  }
  if (timed) HIPCHK(hipEventRecord(ctx->ev1, st));
  ctx->edge_timed = timed && one;
  HIPCHK(hipGetLastError());
  return CDJ_OK;
}

int default_tile(int hs, int64_t nobs) {
  (void)nobs;
  return hs ? 4 : 16;
}

int check_flags(cdj_ctx* ctx, int32_t disp_flag, int32_t strain_flag, double* const* strain, double* const* disp) {
  if (strain_flag == 1 && !strain) return fail(ctx, CDJ_ERR_ARG, "strain outputs missing");
  if (disp_flag == 1 && !disp) return fail(ctx, CDJ_ERR_ARG, "displacement outputs missing");
  if (strain_flag == 1)
    for (int k = 0; k < 18; ++k)
      if (!strain[k]) return fail(ctx, CDJ_ERR_ARG, "strain output pointer is NULL");
  if (disp_flag == 1)
    for (int k = 0; k < 9; ++k)
      if (!disp[k]) return fail(ctx, CDJ_ERR_ARG, "displacement output pointer is NULL");
  return CDJ_OK;
}

// Half-space validity check on device arrays (TD.jl:606-619); synchronises the stream.
int check_hs_dev(cdj_ctx* ctx, const double* Z, int64_t nobs, const double* P1, const double* P2, const double* P3,
                 int64_t ntri, hipStream_t st) {
  DBuf flag;
  HIPCHK(flag.alloc(sizeof(int)));
  HIPCHK(hipMemsetAsync(flag.p, 0, sizeof(int), st));
  launch_check_hs(Z, nobs, P1, P2, P3, ntri, flag.as<int>(), st);
  HIPCHK(hipGetLastError());
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (h) return fail(ctx, CDJ_ERR_HALFSPACE, "Half-space solution: Z coordinates must be negative!");
  return CDJ_OK;
}

// FP32-far / FP64-near pass pair into `out` (the assembly's mixed mode).  rec: FP64 records (built), rec32
// (RS * n floats), near (NEAR_REC n doubles) and cnt (one counter) are device scratch of the caller.
hipError_t traction_mixed(const double* recv, const double* rec, float* rec32, double* near,
                          unsigned long long* cnt, const double* P1, const double* P2, const double* P3, int64_t n,
                          int hs, int tpb, double nu, double lambda, double G, double near_factor,
                          const OutPtrs<double>& out, hipEvent_t ev0, hipEvent_t ev1, hipStream_t st) {
  const int RS = hs ? REC_HS : REC_FS;
  launch_rec_to_f32(rec, rec32, (int64_t)RS * n, st);
  launch_near_info(P1, P2, P3, n, near_factor, near, st);
  launch_zero_u32(reinterpret_cast<unsigned*>(cnt), 2, st);
  hipError_t e = hipSuccess;
  if (ev0 && (e = hipEventRecord(ev0, st)) != hipSuccess) return e;
  launch_traction_f32_to_f64(recv, n, rec32, n, hs, hs ? 8 : 32, (float)nu, (float)lambda, (float)G, out, st);
  launch_nearfield(recv, n, rec, near, n, hs, tpb, nu, lambda, G, out, cnt, st);
  if (ev1 && (e = hipEventRecord(ev1, st)) != hipSuccess) return e;
  return hipGetLastError();
}

int assembly_mode(cdj_ctx* ctx, const cdj_opts* opts, bool* mixed, double* near_factor) {
  *mixed = opts && opts->assembly == CDJ_ASM_MIXED;
  if (opts && opts->assembly != CDJ_ASM_FP64 && !*mixed) return fail(ctx, CDJ_ERR_ARG, "unknown assembly mode");
  *near_factor = (opts && opts->near_factor > 0) ? opts->near_factor : 2.0;
  return CDJ_OK;
}

int64_t count_fixed(const int32_t* fixed, int64_t n) {
  int64_t c = 0;
  if (fixed)
    for (int64_t i = 0; i < n; ++i) c += fixed[i] == 1;
  return c;
}

// Fixed-element system from host geometry (CalculateInfluenceMatrices3D.jl:10-33, 145-198), in FP64 whatever the
// assembly option: block k of `out` (device, n rows x nf columns) receives
//   * at a free element's row, the traction of the strain at its collocation point.  The reference hands only the
//     free strain rows to ConvertStrainInfMatsToTraction with the full normal arrays, so the i-th free row takes
//     the plane (normal, strike and dip directions) of element i (CalculateInfluenceMatrices3D.jl:24-26, 73-110);
//   * at a fixed element's row, the displacement of its collocation point (DnTn <- DnUx, DnTss <- DnUy, ...).
// Only free sources are evaluated: the reference drops the fixed sources' columns (PutDispInfsIntoInfMat).
int assemble_fixed(cdj_ctx* ctx, const double* normals, const double* midpoints, const double* P1, const double* P2,
                   const double* P3, int64_t n, const int32_t* fixed, double nu, double G, double lambda, int hs,
                   const cdj_opts* opts, const OutPtrs<double>& out, hipStream_t st) {
  std::vector<int64_t> fr, fx;
  for (int64_t i = 0; i < n; ++i) (fixed[i] == 1 ? fx : fr).push_back(i);
  const int64_t nf = (int64_t)fr.size(), nfix = (int64_t)fx.size();
  // collocation points (SetupCollationPoints, CalculateInfluenceMatrices3D.jl:52-71)
  std::vector<double> pts(3 * n);
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) pts[k * n + i] = midpoints[k * n + i] - (normals[k * n + i] * 1e-12);
  if (hs) {
    for (int64_t i = 0; i < n; ++i)
      if (pts[2 * n + i] > 0 || P1[2 * n + i] > 0 || P2[2 * n + i] > 0 || P3[2 * n + i] > 0)
        return fail(ctx, CDJ_ERR_HALFSPACE, "Half-space solution: Z coordinates must be negative!");
  }
  if (nf == 0) return CDJ_OK;
  // host gathers: free receivers (points, planes of elements 0..nf-1), fixed receivers, free sources
  std::vector<double> pf(3 * nf), nr(3 * nf), px(3 * std::max<int64_t>(nfix, 1)), s1(3 * nf), s2(3 * nf),
      s3(3 * nf), ones(nf, 1.0);
  std::vector<int32_t> rowmap(n);
  for (int64_t t = 0; t < nf; ++t) {
    const int64_t i = fr[t];
    rowmap[i] = (int32_t)t;
    for (int k = 0; k < 3; ++k) {
      pf[k * nf + t] = pts[k * n + i];
      nr[k * nf + t] = normals[k * n + t];
      s1[k * nf + t] = P1[k * n + i];
      s2[k * nf + t] = P2[k * n + i];
      s3[k * nf + t] = P3[k * n + i];
    }
  }
  for (int64_t t = 0; t < nfix; ++t) {
    rowmap[fx[t]] = (int32_t)(-1 - t);
    for (int k = 0; k < 3; ++k) px[k * nfix + t] = pts[k * n + fx[t]];
  }
  const int RS = hs ? REC_HS : REC_FS;
  const size_t v3 = sizeof(double) * 3 * nf;
  DBuf dpf, dnr, ds1, ds2, ds3, dpx, done, dmap, recv, rec, Tf, Ud;
  HIPCHK(dpf.alloc(v3));
  HIPCHK(dnr.alloc(v3));
  HIPCHK(ds1.alloc(v3));
  HIPCHK(ds2.alloc(v3));
  HIPCHK(ds3.alloc(v3));
  HIPCHK(dpx.alloc(sizeof(double) * px.size()));
  HIPCHK(done.alloc(sizeof(double) * nf));
  HIPCHK(dmap.alloc(sizeof(int32_t) * n));
  HIPCHK(recv.alloc(sizeof(double) * RECV * nf));
  HIPCHK(rec.alloc(sizeof(double) * RS * nf));
  HIPCHK(Tf.alloc(sizeof(double) * 9 * nf * nf));
  HIPCHK(Ud.alloc(sizeof(double) * 9 * std::max<int64_t>(nfix, 1) * nf));
  HIPCHK(hipMemcpyAsync(dpf.p, pf.data(), v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dnr.p, nr.data(), v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ds1.p, s1.data(), v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ds2.p, s2.data(), v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ds3.p, s3.data(), v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dpx.p, px.data(), sizeof(double) * px.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(done.p, ones.data(), sizeof(double) * nf, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dmap.p, rowmap.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  launch_receivers_at(dnr.as<double>(), dpf.as<double>(), nf, recv.as<double>(), st);
  launch_setup(ds1.as<double>(), ds2.as<double>(), ds3.as<double>(), nf, hs, rec.as<double>(), st);
  OutPtrs<double> ot{};
  ot.ld = nf;
  ot.sign = 1.0;
  for (int k = 0; k < 9; ++k) ot.p[k] = Tf.as<double>() + (int64_t)k * nf * nf;
  const int tpb = (opts && opts->tile > 0) ? opts->tile : default_tile(hs, nf);
  launch_traction<double>(recv.as<double>(), nf, rec.as<double>(), nf, hs, tpb, nu, lambda, G, ot, st);
  if (nfix > 0) {
    OutPtrs<double> od{};
    od.ld = nfix;
    od.sign = 1.0;
    od.disp_on = 1;
    for (int k = 0; k < 9; ++k) od.p[18 + k] = Ud.as<double>() + (int64_t)k * nfix * nf;
    const double* x = dpx.as<double>();
    launch_influence<double>(x, x + nfix, x + 2 * nfix, nfix, rec.as<double>(), nf, hs, default_tile(hs, nfix),
                             done.as<double>(), done.as<double>(), done.as<double>(), nu, od, st);
  }
  launch_fixed_scatter(Tf.as<double>(), Ud.as<double>(), dmap.as<int32_t>(), n, nf, nfix, nf, out, st);
  HIPCHK(hipGetLastError());
  // the scoped buffers are released on return, so finish with them first
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

}  // namespace

extern "C" {

int cdj_create(int device, cdj_ctx** out) {
  if (!out) return CDJ_ERR_ARG;
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    g_create_err = std::string("no HIP device: ") + hipGetErrorString(e);
    return CDJ_ERR_DEVICE;
  }
  if (device < 0 || device >= count) {
    g_create_err = "device index out of range";
    return CDJ_ERR_ARG;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_create_err = hipGetErrorString(e);
    return CDJ_ERR_DEVICE;
  }
  cdj_ctx* c = new cdj_ctx();
  c->device = device;
  e = noalloc_scratch_ptr(device, &c->noalloc_scratch);
  if (e != hipSuccess) {
    g_create_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    delete c;
    return CDJ_ERR_DEVICE;
  }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess) {
    g_create_err = "hipEventCreate failed";
    delete c;
    return CDJ_ERR_DEVICE;
  }
  // A blocking stream: it is ordered with the legacy null stream, which is where a caller that passes stream 0
  // (a framework's default stream) has queued the work producing our inputs and consuming our outputs.
  e = hipStreamCreateWithFlags(&c->stream, hipStreamDefault);
  // the LU's look-ahead stream: non-blocking (ordered with the caller's stream only through events)
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_create_err = hipGetErrorString(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CDJ_ERR_DEVICE;
  }
  *out = c;
  return CDJ_OK;
}

int cdj_destroy(cdj_ctx* ctx) {
  if (!ctx) return CDJ_OK;
  (void)hipSetDevice(ctx->device);
  // drain the stream, then release what was made on it (graphs and their buffers) before the stream itself
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  batch_plan_free(ctx->batch);
  ctx->batch = nullptr;
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->scratch_topo) (void)hipFree(ctx->scratch_topo);
  if (ctx->scratch_resp) (void)hipFree(ctx->scratch_resp);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
  if (ctx->ev3) (void)hipEventDestroy(ctx->ev3);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CDJ_OK;
}

const char* cdj_last_error(const cdj_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

double cdj_last_kernel_ms(cdj_ctx* ctx) {
  if (!ctx || !ctx->timed) return -1.0;
  float ms = -1.f;
  if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) return -1.0;
  return (double)ms;
}

int64_t cdj_last_near_entries(const cdj_ctx* ctx) { return ctx ? ctx->near_entries : -1; }

int cdj_last_edge_stats(cdj_ctx* ctx, int64_t* edges, double* ms) {
  if (!ctx) return CDJ_ERR_ARG;
  if (edges) *edges = ctx->last_edges;
  if (ms) {
    float t = -1.f;
    if (!ctx->edge_timed || hipEventElapsedTime(&t, ctx->ev2, ctx->ev3) != hipSuccess) t = -1.f;
    *ms = (double)t;
  }
  return CDJ_OK;
}

int cdj_td_influence_dev(cdj_ctx* ctx, const double* X, const double* Y, const double* Z, int64_t nobs,
                         const double* P1, const double* P2, const double* P3, int64_t ntri, const double* Dss,
                         const double* Dds, const double* Dn, double nu, double G, int32_t disp_flag,
                         int32_t strain_flag, int32_t hs_flag, double* const* strain, double* const* disp,
                         void* stream) {
  (void)G;
  if (!ctx) return CDJ_ERR_ARG;
  if (nobs < 0 || ntri < 0) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  int rc = check_flags(ctx, disp_flag, strain_flag, strain, disp);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  HIPCHK(hipSetDevice(ctx->device));
  const int hs = hs_flag == 1;
  if (hs) {
    rc = check_hs_dev(ctx, Z, nobs, P1, P2, P3, ntri, st);
    if (rc) return rc;
  }
  if (nobs == 0 || ntri == 0 || (disp_flag != 1 && strain_flag != 1)) return CDJ_OK;
  const int RS = hs ? REC_HS : REC_FS;
  DBuf rec;
  HIPCHK(rec.alloc(sizeof(double) * RS * ntri));
  launch_setup(P1, P2, P3, ntri, hs, rec.as<double>(), st);
  OutPtrs<double> out{};
  out.ld = nobs;
  out.sign = 1.0;
  out.strain_on = strain_flag == 1;
  out.disp_on = disp_flag == 1;
  if (out.strain_on)
    for (int k = 0; k < 18; ++k) out.p[k] = strain[k];
  if (out.disp_on)
    for (int k = 0; k < 9; ++k) out.p[18 + k] = disp[k];
  launch_influence<double>(X, Y, Z, nobs, rec.as<double>(), ntri, hs, default_tile(hs, nobs), Dn, Dss, Dds, nu, out,
                           st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

int cdj_td_influence(cdj_ctx* ctx, const double* X, const double* Y, const double* Z, int64_t nobs,
                     const double* P1, const double* P2, const double* P3, int64_t ntri, const double* Dss,
                     const double* Dds, const double* Dn, double nu, double G, int32_t disp_flag,
                     int32_t strain_flag, int32_t hs_flag, double* const* strain, double* const* disp) {
  if (!ctx) return CDJ_ERR_ARG;
  if (nobs < 0 || ntri < 0) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  int rc = check_flags(ctx, disp_flag, strain_flag, strain, disp);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t vo = sizeof(double) * nobs, vt = sizeof(double) * ntri;
  DBuf dX, dY, dZ, dP1, dP2, dP3, dDn, dDss, dDds;
  HIPCHK(dX.alloc(vo));
  HIPCHK(dY.alloc(vo));
  HIPCHK(dZ.alloc(vo));
  HIPCHK(dP1.alloc(3 * vt));
  HIPCHK(dP2.alloc(3 * vt));
  HIPCHK(dP3.alloc(3 * vt));
  HIPCHK(dDn.alloc(vt));
  HIPCHK(dDss.alloc(vt));
  HIPCHK(dDds.alloc(vt));
  HIPCHK(hipMemcpyAsync(dX.p, X, vo, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dY.p, Y, vo, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dZ.p, Z, vo, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dP1.p, P1, 3 * vt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dP2.p, P2, 3 * vt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dP3.p, P3, 3 * vt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dDn.p, Dn, vt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dDss.p, Dss, vt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dDds.p, Dds, vt, hipMemcpyHostToDevice, st));
  const size_t vm = sizeof(double) * nobs * ntri;
  const int nout = (strain_flag == 1 ? 18 : 0) + (disp_flag == 1 ? 9 : 0);
  DBuf dout;
  HIPCHK(dout.alloc(vm * nout));
  double* ds[18] = {};
  double* dd[9] = {};
  int q = 0;
  if (strain_flag == 1)
    for (int k = 0; k < 18; ++k) ds[k] = dout.as<double>() + (q++) * nobs * ntri;
  if (disp_flag == 1)
    for (int k = 0; k < 9; ++k) dd[k] = dout.as<double>() + (q++) * nobs * ntri;
  rc = cdj_td_influence_dev(ctx, dX.as<double>(), dY.as<double>(), dZ.as<double>(), nobs, dP1.as<double>(),
                            dP2.as<double>(), dP3.as<double>(), ntri, dDss.as<double>(), dDds.as<double>(),
                            dDn.as<double>(), nu, G, disp_flag, strain_flag, hs_flag, strain_flag == 1 ? ds : nullptr,
                            disp_flag == 1 ? dd : nullptr, st);
  if (rc) return rc;
  if (strain_flag == 1)
    for (int k = 0; k < 18; ++k) HIPCHK(hipMemcpyAsync(strain[k], ds[k], vm, hipMemcpyDeviceToHost, st));
  if (disp_flag == 1)
    for (int k = 0; k < 9; ++k) HIPCHK(hipMemcpyAsync(disp[k], dd[k], vm, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

int cdj_td_field(cdj_ctx* ctx, const double* X, const double* Y, const double* Z, int64_t nobs, const double* P1,
                 const double* P2, const double* P3, int64_t ntri, const double* Dss, const double* Dds,
                 const double* Dn, double nu, double G, int32_t disp_flag, int32_t strain_flag, int32_t hs_flag,
                 double* const* strain6, double* const* disp3) {
  (void)G;
  if (!ctx) return CDJ_ERR_ARG;
  if (nobs < 0 || ntri < 0) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  if (strain_flag == 1 && !strain6) return fail(ctx, CDJ_ERR_ARG, "strain outputs missing");
  if (disp_flag == 1 && !disp3) return fail(ctx, CDJ_ERR_ARG, "displacement outputs missing");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int hs = hs_flag == 1;
  const size_t vo = sizeof(double) * nobs, vt = sizeof(double) * ntri;
  DBuf dX, dY, dZ, dP1, dP2, dP3, dDn, dDss, dDds;
  HIPCHK(dX.alloc(vo));
  HIPCHK(dY.alloc(vo));
  HIPCHK(dZ.alloc(vo));
  HIPCHK(dP1.alloc(3 * vt));
  HIPCHK(dP2.alloc(3 * vt));
  HIPCHK(dP3.alloc(3 * vt));
  HIPCHK(dDn.alloc(vt));
  HIPCHK(dDss.alloc(vt));
  HIPCHK(dDds.alloc(vt));
  HIPCHK(hipMemcpyAsync(dX.p, X, vo, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dY.p, Y, vo, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dZ.p, Z, vo, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dP1.p, P1, 3 * vt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dP2.p, P2, 3 * vt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dP3.p, P3, 3 * vt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dDn.p, Dn, vt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dDss.p, Dss, vt, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dDds.p, Dds, vt, hipMemcpyHostToDevice, st));
  if (hs) {
    int rc = check_hs_dev(ctx, dZ.as<double>(), nobs, dP1.as<double>(), dP2.as<double>(), dP3.as<double>(), ntri, st);
    if (rc) return rc;
  }
  const int strain_on = strain_flag == 1, disp_on = disp_flag == 1;
  const int ncomp = (strain_on ? 6 : 0) + (disp_on ? 3 : 0);
  if (nobs == 0 || ncomp == 0) return CDJ_OK;
  const int tpb = hs ? 8 : 32;
  const int64_t nchunk = ntri > 0 ? (ntri + tpb - 1) / tpb : 0;
  const int RS = hs ? REC_HS : REC_FS;
  DBuf rec, part, dout;
  HIPCHK(rec.alloc(sizeof(double) * RS * std::max<int64_t>(ntri, 1)));
  HIPCHK(part.alloc(sizeof(double) * std::max<int64_t>(nchunk, 1) * ncomp * nobs));
  HIPCHK(dout.alloc(sizeof(double) * ncomp * nobs));
  launch_setup(dP1.as<double>(), dP2.as<double>(), dP3.as<double>(), ntri, hs, rec.as<double>(), st);
  launch_field<double>(dX.as<double>(), dY.as<double>(), dZ.as<double>(), nobs, rec.as<double>(), ntri, hs, tpb,
                       dDn.as<double>(), dDss.as<double>(), dDds.as<double>(), nu, strain_on, disp_on,
                       part.as<double>(), dout.as<double>(), st);
  HIPCHK(hipGetLastError());
  int q = 0;
  if (strain_on)
    for (int k = 0; k < 6; ++k)
      HIPCHK(hipMemcpyAsync(strain6[k], dout.as<double>() + (q++) * nobs, vo, hipMemcpyDeviceToHost, st));
  if (disp_on)
    for (int k = 0; k < 3; ++k)
      HIPCHK(hipMemcpyAsync(disp3[k], dout.as<double>() + (q++) * nobs, vo, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

}  // extern "C"

namespace {

// Device scratch of the traction assembly of n triangles: a 256-byte header (the near-entry counter at 0, the
// half-space check flag at 8), then receiver records, triangle records and, for the mixed mode, FP32 records and
// the near-field table.
size_t assemble_scratch_bytes(int64_t n, int hs, bool mixed) {
  const size_t RS = hs ? REC_HS : REC_FS;
  size_t b = 256 + sizeof(double) * (RECV + RS) * (size_t)n;
  if (mixed) b += ((sizeof(float) * RS * (size_t)n + 255) / 256) * 256 + sizeof(double) * NEAR_REC * (size_t)n;
  return b;
}

// cdj_assemble_dev on the scratch `scr` (assemble_scratch_bytes(n, hs, mixed) bytes of device memory, 256-byte
// aligned).  With share_edges the FP64 half-space assembly evaluates each mesh edge's free-surface part once
// (traction_hs_fp64, which keeps its own context buffers); without it nothing is allocated and every pair
// evaluates its triangle's three sides.
int assemble_into(cdj_ctx* ctx, const double* normals, const double* midpoints, const double* P1, const double* P2,
                  const double* P3, int64_t n, double nu, double G, double lambda, int hs, bool mixed,
                  double near_factor, double* A, int64_t lda, const cdj_opts* opts, char* scr, bool share_edges,
                  hipStream_t st) {
  const int RS = hs ? REC_HS : REC_FS;
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(scr);
  int* flag = reinterpret_cast<int*>(scr + 8);
  double* recv_p = reinterpret_cast<double*>(scr + 256);
  double* rec_p = recv_p + RECV * n;
  float* rec32 = reinterpret_cast<float*>(rec_p + RS * n);
  double* near_p = reinterpret_cast<double*>(reinterpret_cast<char*>(rec32) +
                                             ((sizeof(float) * RS * (size_t)n + 255) / 256) * 256);
  launch_receivers(normals, midpoints, n, recv_p, st);
  if (hs) {
    // collocation points (z in the receiver records) and triangle corners must lie in z <= 0 (TD.jl:606-619)
    launch_zero_u32(reinterpret_cast<unsigned*>(flag), 1, st);
    launch_check_hs(recv_p + 2, n, P1, P2, P3, n, flag, st, RECV);
    HIPCHK(hipGetLastError());
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h) return fail(ctx, CDJ_ERR_HALFSPACE, "Half-space solution: Z coordinates must be negative!");
  }
  launch_setup(P1, P2, P3, n, hs, rec_p, st);
  OutPtrs<double> out{};
  out.ld = lda;
  out.sign = -1.0;
  for (int m = 0; m < 3; ++m)
    for (int q = 0; q < 3; ++q) out.p[3 * m + q] = A + (int64_t)q * n + (int64_t)m * n * lda;
  const int tpb = (opts && opts->tile > 0) ? opts->tile : default_tile(hs, n);
  ctx->edge_timed = false;
  ctx->last_edges = 0;
  if (!mixed) {
    if (hs && share_edges) {
      if (int rc = traction_hs_fp64(ctx, recv_p, n, P1, P2, P3, rec_p, n, tpb, nu, lambda, G, out, true, st)) return rc;
    } else {
      HIPCHK(hipEventRecord(ctx->ev0, st));
      launch_traction<double>(recv_p, n, rec_p, n, hs, tpb, nu, lambda, G, out, st);
      HIPCHK(hipEventRecord(ctx->ev1, st));
    }
    ctx->timed = true;
    ctx->near_entries = 0;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return CDJ_OK;
  }
  // the timed region covers both passes (the mixed assembly's cost)
  HIPCHK(traction_mixed(recv_p, rec_p, rec32, near_p, cnt, P1, P2, P3, n, hs, tpb, nu, lambda, G, near_factor, out,
                        ctx->ev0, ctx->ev1, st));
  ctx->timed = true;
  unsigned long long h = 0;
  HIPCHK(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  ctx->near_entries = (int64_t)h;
  return CDJ_OK;
}

}  // namespace

extern "C" {

int cdj_assemble_dev(cdj_ctx* ctx, const double* normals, const double* midpoints, const double* P1,
                     const double* P2, const double* P3, int64_t n, double nu, double G, double lambda,
                     int32_t hs_flag, double* A, int64_t lda, const cdj_opts* opts, void* stream) {
  if (!ctx) return CDJ_ERR_ARG;
  if (n < 0 || lda < 3 * n) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  if (n == 0) return CDJ_OK;
  const int hs = hs_flag == 1;
  bool mixed = false;
  double near_factor = 2.0;
  if (int rc = assembly_mode(ctx, opts, &mixed, &near_factor)) return rc;
  void* scr = nullptr;
  HIPCHK(ctx_scratch(ctx, assemble_scratch_bytes(n, hs, mixed), &scr));
  return assemble_into(ctx, normals, midpoints, P1, P2, P3, n, nu, G, lambda, hs, mixed, near_factor, A, lda, opts,
                       static_cast<char*>(scr), true, st);
}

int cdj_assemble_dev_noalloc(cdj_ctx* ctx, const double* normals, const double* midpoints, const double* P1,
                             const double* P2, const double* P3, int64_t n, double nu, double G, double lambda,
                             int32_t hs_flag, double* A, int64_t lda, const cdj_opts* opts, void* stream) {
  if (!ctx) return CDJ_ERR_ARG;
  if (n < 0 || lda < 3 * n) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  const int hs = hs_flag == 1;
  bool mixed = false;
  double near_factor = 2.0;
  if (int rc = assembly_mode(ctx, opts, &mixed, &near_factor)) return rc;
  const size_t need = assemble_scratch_bytes(n, hs, mixed);
  if (need > CDJ_NOALLOC_SCRATCH_BYTES)
    return fail(ctx, CDJ_ERR_ARG,
                "cdj_assemble_dev_noalloc: " + std::to_string(n) + " triangles need " + std::to_string(need) +
                    " bytes of scratch, above the fixed " + std::to_string((long long)CDJ_NOALLOC_SCRATCH_BYTES) +
                    " bytes (use cdj_assemble_dev)");
  if (!ctx->noalloc_scratch) return fail(ctx, CDJ_ERR_DEVICE, "no-allocation scratch address unavailable");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  if (n == 0) return CDJ_OK;
  return assemble_into(ctx, normals, midpoints, P1, P2, P3, n, nu, G, lambda, hs, mixed, near_factor, A, lda, opts,
                       ctx->noalloc_scratch, false, st);
}

int cdj_traction_influence(cdj_ctx* ctx, const double* normals, const double* midpoints, const double* P1,
                           const double* P2, const double* P3, int64_t n, double nu, double G, double lambda,
                           int32_t hs_flag, const int32_t* fixed, double* const* T, const cdj_opts* opts) {
  if (!ctx) return CDJ_ERR_ARG;
  if (n < 0) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  for (int k = 0; k < 9; ++k)
    if (!T || !T[k]) return fail(ctx, CDJ_ERR_ARG, "traction output pointer is NULL");
  bool mixed = false;
  double near_factor = 2.0;
  if (int rc = assembly_mode(ctx, opts, &mixed, &near_factor)) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (n == 0) return CDJ_OK;
  const int hs = hs_flag == 1;
  if (const int64_t nfix = count_fixed(fixed, n)) {
    if (!normals || !midpoints || !P1 || !P2 || !P3) return fail(ctx, CDJ_ERR_ARG, "NULL argument");
    const int64_t nf = n - nfix;
    const size_t vm = sizeof(double) * n * nf;
    DBuf dT;
    HIPCHK(dT.alloc(9 * vm));
    OutPtrs<double> out{};
    out.ld = n;
    out.sign = 1.0;
    for (int k = 0; k < 9; ++k) out.p[k] = dT.as<double>() + (int64_t)k * n * nf;
    int rc = assemble_fixed(ctx, normals, midpoints, P1, P2, P3, n, fixed, nu, G, lambda, hs, opts, out, st);
    if (rc) return rc;
    ctx->near_entries = 0;
    if (nf > 0)
      for (int k = 0; k < 9; ++k) HIPCHK(hipMemcpyAsync(T[k], out.p[k], vm, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return CDJ_OK;
  }
  const size_t v3 = sizeof(double) * 3 * n;
  DBuf dn, dm, d1, d2, d3, recv, rec, dT;
  HIPCHK(dn.alloc(v3));
  HIPCHK(dm.alloc(v3));
  HIPCHK(d1.alloc(v3));
  HIPCHK(d2.alloc(v3));
  HIPCHK(d3.alloc(v3));
  HIPCHK(hipMemcpyAsync(dn.p, normals, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dm.p, midpoints, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d1.p, P1, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d2.p, P2, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d3.p, P3, v3, hipMemcpyHostToDevice, st));
  HIPCHK(recv.alloc(sizeof(double) * RECV * n));
  const int RS = hs ? REC_HS : REC_FS;
  HIPCHK(rec.alloc(sizeof(double) * RS * n));
  const size_t vm = sizeof(double) * n * n;
  HIPCHK(dT.alloc(9 * vm));
  launch_receivers(dn.as<double>(), dm.as<double>(), n, recv.as<double>(), st);
  if (hs) {
    std::vector<double> rv(RECV * n);
    HIPCHK(hipMemcpyAsync(rv.data(), recv.p, sizeof(double) * RECV * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; ++i)
      if (rv[RECV * i + 2] > 0 || P1[2 * n + i] > 0 || P2[2 * n + i] > 0 || P3[2 * n + i] > 0)
        return fail(ctx, CDJ_ERR_HALFSPACE, "Half-space solution: Z coordinates must be negative!");
  }
  launch_setup(d1.as<double>(), d2.as<double>(), d3.as<double>(), n, hs, rec.as<double>(), st);
  OutPtrs<double> out{};
  out.ld = n;
  out.sign = 1.0;
  for (int k = 0; k < 9; ++k) out.p[k] = dT.as<double>() + (int64_t)k * n * n;
  const int tpb = (opts && opts->tile > 0) ? opts->tile : default_tile(hs, n);
  if (mixed) {
    DBuf r32, nr, cnt;
    HIPCHK(r32.alloc(sizeof(float) * RS * n));
    HIPCHK(nr.alloc(sizeof(double) * NEAR_REC * n));
    HIPCHK(cnt.alloc(sizeof(unsigned long long)));
    HIPCHK(traction_mixed(recv.as<double>(), rec.as<double>(), r32.as<float>(), nr.as<double>(),
                          cnt.as<unsigned long long>(), d1.as<double>(), d2.as<double>(), d3.as<double>(), n, hs, tpb,
                          nu, lambda, G, near_factor, out, nullptr, nullptr, st));
    unsigned long long h = 0;
    HIPCHK(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->near_entries = (int64_t)h;
  } else {
    if (hs) {
      if (int rc = traction_hs_fp64(ctx, recv.as<double>(), n, d1.as<double>(), d2.as<double>(), d3.as<double>(),
                                    rec.as<double>(), n, tpb, nu, lambda, G, out, false, st))
        return rc;
    } else {
      launch_traction<double>(recv.as<double>(), n, rec.as<double>(), n, hs, tpb, nu, lambda, G, out, st);
    }
    ctx->near_entries = 0;
  }
  HIPCHK(hipGetLastError());
  for (int k = 0; k < 9; ++k) HIPCHK(hipMemcpyAsync(T[k], out.p[k], vm, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

}  // extern "C"

int cdj_lu_factor_dev(cdj_ctx* ctx, double* A, int64_t n, int64_t lda, int32_t* ipiv, void* stream) {
  if (!ctx) return CDJ_ERR_ARG;
  if (n < 0 || lda < n) return fail(ctx, CDJ_ERR_DIMS, "invalid matrix dimensions");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  if (n == 0) return CDJ_OK;
  DBuf ws, info;
  HIPCHK(ws.alloc(lu_workspace_bytes(n)));
  HIPCHK(info.alloc(sizeof(int)));
  HIPCHK(lu_factor<double>(A, n, lda, ipiv, info.as<int>(), ws.p, st, false, ctx->side));
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (h < 0) return fail(ctx, CDJ_ERR_DEVICE, "LU panel barrier wait timed out");
  if (h) return fail(ctx, CDJ_ERR_SINGULAR, "matrix is exactly singular (zero pivot " + std::to_string(h) + ")");
  return CDJ_OK;
}

int cdj_lu_factor_batched_dev(cdj_ctx* ctx, double* const* A, const int64_t* n, const int64_t* lda, int64_t count,
                              int32_t* const* ipiv, int32_t* info, void* stream) {
  if (!ctx) return CDJ_ERR_ARG;
  if (count < 0 || (count > 0 && (!A || !n || !lda || !ipiv || !info))) return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  for (int64_t j = 0; j < count; ++j)
    if (n[j] < 0 || lda[j] < n[j] || (n[j] > 0 && (!A[j] || !ipiv[j])))
      return fail(ctx, CDJ_ERR_DIMS, "invalid matrix dimensions");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  std::vector<int64_t> idx;
  size_t ws_total = 0;
  std::vector<size_t> off(count, 0);
  for (int64_t j = 0; j < count; ++j) {
    info[j] = 0;
    if (n[j] == 0) continue;
    idx.push_back(j);
    off[j] = ws_total;
    ws_total += (lu_workspace_bytes(n[j]) + 255) / 256 * 256;
  }
  if (idx.empty()) return CDJ_OK;
  DBuf ws, dinfo;
  HIPCHK(ws.alloc(ws_total));
  HIPCHK(dinfo.alloc(sizeof(int) * idx.size()));
  std::vector<LuJob<double>> jobs;
  for (size_t t = 0; t < idx.size(); ++t) {
    const int64_t j = idx[t];
    jobs.push_back(LuJob<double>{A[j], n[j], n[j], lda[j], ipiv[j], dinfo.as<int>() + t, ws.as<char>() + off[j]});
  }
  HIPCHK(lu_factor_batched<double>(jobs.data(), (int)jobs.size(), st));
  std::vector<int> h(idx.size());
  HIPCHK(hipMemcpyAsync(h.data(), dinfo.p, sizeof(int) * h.size(), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (size_t t = 0; t < idx.size(); ++t) {
    if (h[t] < 0) return fail(ctx, CDJ_ERR_DEVICE, "LU panel barrier wait timed out");
    info[idx[t]] = h[t];
  }
  return CDJ_OK;
}

int cdj_lu_solve_dev(cdj_ctx* ctx, const double* LU, int64_t n, int64_t lda, const int32_t* ipiv, double* B,
                     int64_t nrhs, int64_t ldb, void* stream) {
  if (!ctx) return CDJ_ERR_ARG;
  if (n < 0 || lda < n || ldb < n || nrhs < 0) return fail(ctx, CDJ_ERR_DIMS, "invalid matrix dimensions");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  if (n == 0 || nrhs == 0) return CDJ_OK;
  HIPCHK(lu_solve<double>(LU, n, lda, ipiv, B, nrhs, ldb, st));
  return CDJ_OK;
}

namespace {

// Factor the augmented device matrix [A | B] (n x (n + nrhs), leading dimension lda) in FP64, back-substitute
// and leave X in the B columns.  Aorig (device, may be NULL) is used for the FP64 residual.
int solve_augmented_fp64(cdj_ctx* ctx, double* AB, int64_t n, int64_t lda, int64_t nrhs, const double* Aorig,
                         const double* Borig, cdj_report* report, hipStream_t st) {
  DBuf ws, info, ipiv;
  HIPCHK(ws.alloc(lu_workspace_bytes(n)));
  HIPCHK(info.alloc(sizeof(int)));
  HIPCHK(ipiv.alloc(sizeof(int32_t) * n));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  HIPCHK(lu_factor_cols<double>(AB, n, n + nrhs, lda, ipiv.as<int32_t>(), info.as<int>(), ws.p, st, false,
                                ctx->side));
  HIPCHK(tri_solve_upper<double>(AB, n, lda, AB + n * lda, nrhs, lda, st));
  HIPCHK(hipEventRecord(e1, st));
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (h < 0) return fail(ctx, CDJ_ERR_DEVICE, "LU panel barrier wait timed out");
  if (h) return fail(ctx, CDJ_ERR_SINGULAR, "matrix is exactly singular (zero pivot " + std::to_string(h) + ")");
  if (report) {
    report->path = CDJ_PREC_FP64;
    report->iterations = 0;
    report->solve_ms = ms;
    report->residual = -1.0;
    if (Aorig && Borig) {
      DBuf r;
      HIPCHK(r.alloc(sizeof(double) * n));
      double worst = 0.0;
      for (int64_t q = 0; q < nrhs; ++q) {
        HIPCHK(hipMemcpyAsync(r.p, Borig + q * n, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        gemv<double>(Aorig, n, n, n, AB + (n + q) * lda, r.as<double>(), -1.0, 1.0, st);
        std::vector<double> hr(n), hb(n);
        HIPCHK(hipMemcpyAsync(hr.data(), r.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(hb.data(), Borig + q * n, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        double nr = 0, nb = 0;
        for (int64_t i = 0; i < n; ++i) {
          nr += hr[i] * hr[i];
          nb += hb[i] * hb[i];
        }
        const double rel = nb > 0 ? std::sqrt(nr / nb) : std::sqrt(nr);
        worst = std::max(worst, rel);
      }
      report->residual = worst;
    }
  }
  return CDJ_OK;
}

// Mixed precision: FP32 copy of A factored on MFMA, FP64 FGMRES-IR against the FP64 A (device, n x n, lda).
// B (n x nrhs, ld n) right-hand sides, X (n x nrhs) solutions.  Falls back to the FP64 LU when the
// refinement does not reach the tolerance; report->path says which one produced X.
int solve_mixed32(cdj_ctx* ctx, const double* A, int64_t n, int64_t lda, const double* B, double* X, int64_t nrhs,
                  const cdj_opts* opts, cdj_report* report, hipStream_t st) {
  GmresOpts go;
  if (opts && opts->tol > 0) go.tol = opts->tol;
  if (opts && opts->max_iter > 0) go.max_iter = opts->max_iter;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  DBuf A32, ipiv, perm, ws, info;
  HIPCHK(A32.alloc(sizeof(float) * n * n));
  HIPCHK(ipiv.alloc(sizeof(int32_t) * n));
  HIPCHK(perm.alloc(sizeof(int32_t) * n));
  HIPCHK(ws.alloc(lu_workspace_bytes(n)));
  HIPCHK(info.alloc(sizeof(int)));
  convert_matrix_d2f(A, lda, A32.as<float>(), n, n, n, st);
  const bool fp16 = opts && opts->precision == CDJ_PREC_MIXED16;
  HIPCHK(lu_factor<float>(A32.as<float>(), n, n, ipiv.as<int32_t>(), info.as<int>(), ws.p, st, fp16, ctx->side));
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(lu_permutation(ipiv.as<int32_t>(), n, perm.as<int32_t>(), st));
  HIPCHK(hipStreamSynchronize(st));  // h is read below
  bool ok = h == 0;
  double worst = 0.0;
  int iters = 0;
  std::string why = h < 0   ? std::string("an FP32 LU panel wait timed out")
                    : h > 0 ? "the FP32 factor has a zero pivot (column " + std::to_string(h) + ")"
                            : std::string();
  for (int64_t q = 0; ok && q < nrhs; ++q) {
    GmresResult gr;
    HIPCHK(fgmres_ir<float>(A, n, lda, A32.as<float>(), n, perm.as<int32_t>(), B + q * n, X + q * n, go, gr, st));
    ok = gr.converged && !gr.trsv_timeout;
    worst = std::max(worst, gr.residual);
    iters += gr.iterations;
    if (!ok)
      why = gr.trsv_timeout ? "a triangular-solve wait timed out (right-hand side " + std::to_string(q) + ")"
                            : "FGMRES-IR stopped at residual " + std::to_string(gr.residual) + " after " +
                                  std::to_string(gr.iterations) + " iterations (right-hand side " +
                                  std::to_string(q) + ")";
  }
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ok) {
    if (report) {
      report->path = fp16 ? CDJ_PREC_MIXED16 : CDJ_PREC_MIXED32;
      report->iterations = iters;
      report->residual = worst;
      report->solve_ms = ms;
    }
    return CDJ_OK;
  }
  // fallback: FP64 LU on a copy of A; the reason stays readable through cdj_last_error
  ctx->err = "mixed-precision refinement fell back to the FP64 LU: " + why;
  DBuf AB;
  HIPCHK(AB.alloc(sizeof(double) * n * (n + nrhs)));
  HIPCHK(hipMemcpy2DAsync(AB.p, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                          hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(AB.as<double>() + n * n, B, sizeof(double) * n * nrhs, hipMemcpyDeviceToDevice, st));
  int rc = solve_augmented_fp64(ctx, AB.as<double>(), n, n, nrhs, nullptr, nullptr, nullptr, st);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(X, AB.as<double>() + n * n, sizeof(double) * n * nrhs, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  if (report) {
    report->path = CDJ_PREC_FP64;
    report->iterations = iters;
    report->residual = -1.0;
    report->solve_ms = ms;
  }
  return CDJ_OK;
}

// Least squares min ||A x - b|| for device A (m x k, k <= m, leading dimension lda): the `A\b` of
// SlipCalculator3D.jl:22 when fixed elements make A non-square (Julia: pivoted QR).  Solved through the augmented
// system [[I, A], [A^T, 0]] [r; x] = [b; 0] with the FP64 LU, so no normal equations are formed.  B (device,
// m x nrhs, ld m) -> X (device, k x nrhs, ld k).  report->residual is the largest ||b - A x|| / ||b|| over the
// right-hand sides: the least-squares residual, which is not small in general.
int lstsq_dev(cdj_ctx* ctx, const double* A, int64_t m, int64_t k, int64_t lda, const double* B, int64_t nrhs,
              double* X, cdj_report* report, hipStream_t st) {
  if (m + k > 65535) return fail(ctx, CDJ_ERR_DIMS, "least-squares system too large (m + k > 65535)");
  if (k == 0 || nrhs == 0) return CDJ_OK;
  const int64_t N = m + k;
  DBuf KB;
  HIPCHK(KB.alloc(sizeof(double) * N * (N + nrhs)));
  build_kkt(A, lda, m, k, KB.as<double>(), N, st);
  double* R = KB.as<double>() + N * N;
  HIPCHK(hipMemsetAsync(R, 0, sizeof(double) * N * nrhs, st));
  HIPCHK(hipMemcpy2DAsync(R, sizeof(double) * N, B, sizeof(double) * m, sizeof(double) * m, nrhs,
                          hipMemcpyDeviceToDevice, st));
  int rc = solve_augmented_fp64(ctx, KB.as<double>(), N, N, nrhs, nullptr, nullptr, report, st);
  if (rc) return rc;
  HIPCHK(hipMemcpy2DAsync(X, sizeof(double) * k, R + m, sizeof(double) * N, sizeof(double) * k, nrhs,
                          hipMemcpyDeviceToDevice, st));
  if (report) {
    // r = b - A x in the B columns of the (no longer needed) augmented right-hand sides
    double worst = 0.0;
    std::vector<double> hr(m), hb(m);
    for (int64_t q = 0; q < nrhs; ++q) {
      double* r = R + q * N;
      HIPCHK(hipMemcpyAsync(r, B + q * m, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
      gemv<double>(A, m, k, lda, X + q * k, r, -1.0, 1.0, st);
      HIPCHK(hipMemcpyAsync(hr.data(), r, sizeof(double) * m, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(hb.data(), B + q * m, sizeof(double) * m, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      double nr = 0, nb = 0;
      for (int64_t i = 0; i < m; ++i) {
        nr += hr[i] * hr[i];
        nb += hb[i] * hb[i];
      }
      worst = std::max(worst, nb > 0 ? std::sqrt(nr / nb) : std::sqrt(nr));
    }
    report->residual = worst;
  }
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

}  // namespace

extern "C" {

int cdj_solve(cdj_ctx* ctx, const double* A, const double* b, int64_t n, int64_t nrhs, double* x,
              const cdj_opts* opts, cdj_report* report) {
  if (!ctx) return CDJ_ERR_ARG;
  if (n < 0 || nrhs < 0) return fail(ctx, CDJ_ERR_DIMS, "invalid matrix dimensions");
  if (!A || !b || !x) return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  (void)opts;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (n == 0 || nrhs == 0) return CDJ_OK;
  const int64_t lda = n;
  if (opts && (opts->precision == CDJ_PREC_MIXED32 || opts->precision == CDJ_PREC_MIXED16)) {
    DBuf dA, dB, dX;
    HIPCHK(dA.alloc(sizeof(double) * n * n));
    HIPCHK(dB.alloc(sizeof(double) * n * nrhs));
    HIPCHK(dX.alloc(sizeof(double) * n * nrhs));
    HIPCHK(hipMemcpyAsync(dA.p, A, sizeof(double) * n * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dB.p, b, sizeof(double) * n * nrhs, hipMemcpyHostToDevice, st));
    int rc = solve_mixed32(ctx, dA.as<double>(), n, n, dB.as<double>(), dX.as<double>(), nrhs, opts, report, st);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(x, dX.p, sizeof(double) * n * nrhs, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return CDJ_OK;
  }
  DBuf AB, Ao, Bo;
  HIPCHK(AB.alloc(sizeof(double) * lda * (n + nrhs)));
  HIPCHK(hipMemcpyAsync(AB.p, A, sizeof(double) * n * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(AB.as<double>() + n * lda, b, sizeof(double) * n * nrhs, hipMemcpyHostToDevice, st));
  if (report) {
    HIPCHK(Ao.alloc(sizeof(double) * n * n));
    HIPCHK(Bo.alloc(sizeof(double) * n * nrhs));
    HIPCHK(hipMemcpyAsync(Ao.p, AB.p, sizeof(double) * n * n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(Bo.p, AB.as<double>() + n * lda, sizeof(double) * n * nrhs, hipMemcpyDeviceToDevice, st));
  }
  int rc = solve_augmented_fp64(ctx, AB.as<double>(), n, lda, nrhs, Ao.as<double>(), Bo.as<double>(), report, st);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(x, AB.as<double>() + n * lda, sizeof(double) * n * nrhs, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

int cdj_slip_calculator(cdj_ctx* ctx, const double* P1, const double* P2, const double* P3, int64_t n, double nu,
                        double G, double lambda, const double* midpoints, const double* normals, int32_t hs_flag,
                        const double* b, int64_t nrhs, const int32_t* fixed, double* x, double* A_out,
                        const cdj_opts* opts, cdj_report* report) {
  if (!ctx) return CDJ_ERR_ARG;
  if (n < 0 || nrhs < 0) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  if (!P1 || !P2 || !P3 || !midpoints || !normals || !b || !x) return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (n == 0 || nrhs == 0) return CDJ_OK;
  if (const int64_t nfix = count_fixed(fixed, n)) {
    // fixed elements: A is 3n x 3nf (displacement rows at fixed elements, fixed sources' columns dropped) and
    // x its least-squares solution (SlipCalculator3D.jl:13-29); FP64 whatever opts->precision
    const int64_t nf = n - nfix, m = 3 * n, k = 3 * nf;
    if (m + k > 65535) return fail(ctx, CDJ_ERR_DIMS, "least-squares system too large (3n + 3nf > 65535)");
    DBuf dA, dB, dX;
    HIPCHK(dA.alloc(sizeof(double) * m * std::max<int64_t>(k, 1)));
    HIPCHK(dB.alloc(sizeof(double) * m * nrhs));
    HIPCHK(dX.alloc(sizeof(double) * std::max<int64_t>(k, 1) * nrhs));
    OutPtrs<double> out{};
    out.ld = m;
    out.sign = -1.0;
    for (int mo = 0; mo < 3; ++mo)
      for (int q = 0; q < 3; ++q) out.p[3 * mo + q] = dA.as<double>() + (int64_t)q * n + (int64_t)mo * nf * m;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    int rc = assemble_fixed(ctx, normals, midpoints, P1, P2, P3, n, fixed, nu, G, lambda, hs_flag == 1, opts, out, st);
    (void)hipEventRecord(e1, st);
    (void)hipStreamSynchronize(st);
    float ams = 0.f;
    (void)hipEventElapsedTime(&ams, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (A_out && k > 0) HIPCHK(hipMemcpyAsync(A_out, dA.p, sizeof(double) * m * k, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(dB.p, b, sizeof(double) * m * nrhs, hipMemcpyHostToDevice, st));
    rc = lstsq_dev(ctx, dA.as<double>(), m, k, m, dB.as<double>(), nrhs, dX.as<double>(), report, st);
    if (rc) return rc;
    if (report) {
      report->assembly_ms = ams;
      report->near_entries = 0;
    }
    ctx->near_entries = 0;
    if (k > 0) HIPCHK(hipMemcpyAsync(x, dX.p, sizeof(double) * k * nrhs, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return CDJ_OK;
  }
  const int64_t N3 = 3 * n, lda = N3;
  const size_t v3 = sizeof(double) * 3 * n;
  DBuf dn, dm, d1, d2, d3, AB, Ao, Bo;
  HIPCHK(dn.alloc(v3));
  HIPCHK(dm.alloc(v3));
  HIPCHK(d1.alloc(v3));
  HIPCHK(d2.alloc(v3));
  HIPCHK(d3.alloc(v3));
  HIPCHK(hipMemcpyAsync(dn.p, normals, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dm.p, midpoints, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d1.p, P1, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d2.p, P2, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d3.p, P3, v3, hipMemcpyHostToDevice, st));
  HIPCHK(AB.alloc(sizeof(double) * lda * (N3 + nrhs)));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  int rc = cdj_assemble_dev(ctx, dn.as<double>(), dm.as<double>(), d1.as<double>(), d2.as<double>(), d3.as<double>(),
                            n, nu, G, lambda, hs_flag, AB.as<double>(), lda, opts, st);
  if (rc) return rc;
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipMemcpyAsync(AB.as<double>() + N3 * lda, b, sizeof(double) * N3 * nrhs, hipMemcpyHostToDevice, st));
  if (A_out) HIPCHK(hipMemcpyAsync(A_out, AB.p, sizeof(double) * N3 * N3, hipMemcpyDeviceToHost, st));
  const bool mixed = opts && (opts->precision == CDJ_PREC_MIXED32 || opts->precision == CDJ_PREC_MIXED16);
  if (report && !mixed) {
    HIPCHK(Ao.alloc(sizeof(double) * N3 * N3));
    HIPCHK(Bo.alloc(sizeof(double) * N3 * nrhs));
    HIPCHK(hipMemcpyAsync(Ao.p, AB.p, sizeof(double) * N3 * N3, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(Bo.p, AB.as<double>() + N3 * lda, sizeof(double) * N3 * nrhs, hipMemcpyDeviceToDevice, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  float ams = 0.f;
  (void)hipEventElapsedTime(&ams, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (mixed) {
    DBuf dX;
    HIPCHK(dX.alloc(sizeof(double) * N3 * nrhs));
    rc = solve_mixed32(ctx, AB.as<double>(), N3, lda, AB.as<double>() + N3 * lda, dX.as<double>(), nrhs, opts, report,
                       st);
    if (rc) return rc;
    if (report) {
      report->assembly_ms = ams;
      report->near_entries = ctx->near_entries;
    }
    HIPCHK(hipMemcpyAsync(x, dX.p, sizeof(double) * N3 * nrhs, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return CDJ_OK;
  }
  rc = solve_augmented_fp64(ctx, AB.as<double>(), N3, lda, nrhs, Ao.as<double>(), Bo.as<double>(), report, st);
  if (rc) return rc;
  if (report) {
    report->assembly_ms = ams;
    report->near_entries = ctx->near_entries;
  }
  HIPCHK(hipMemcpyAsync(x, AB.as<double>() + N3 * lda, sizeof(double) * N3 * nrhs, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

}  // extern "C"

// ---- learned preconditioner -----------------------------------------------------------------------------
extern "C" {

int cdj_mlp_create(cdj_ctx* ctx, int32_t k, int32_t hidden, uint64_t seed, cdj_mlp** out) {
  if (!ctx || !out) return CDJ_ERR_ARG;
  if (k < 1 || k > KNN_MAX || hidden < 1) return fail(ctx, CDJ_ERR_ARG, "invalid MLP shape");
  HIPCHK(hipSetDevice(ctx->device));
  cdj_mlp* p = new cdj_mlp();
  hipError_t e = mlp_init(p->m, k, hidden, seed);
  if (e != hipSuccess) {
    mlp_free(p->m);
    delete p;
    return hip_fail(ctx, e, "mlp_init");
  }
  *out = p;
  return CDJ_OK;
}

int cdj_mlp_destroy(cdj_mlp* mlp) {
  if (mlp) {
    mlp_free(mlp->m);
    delete mlp;
  }
  return CDJ_OK;
}

int64_t cdj_mlp_nparams(const cdj_mlp* mlp) { return mlp ? mlp->m.nparams() : 0; }

int cdj_mlp_get_params(cdj_ctx* ctx, const cdj_mlp* mlp, float* params) {
  if (!ctx || !mlp || !params) return CDJ_ERR_ARG;
  HIPCHK(hipMemcpy(params, mlp->m.params, sizeof(float) * mlp->m.nparams(), hipMemcpyDeviceToHost));
  return CDJ_OK;
}

int cdj_mlp_set_params(cdj_ctx* ctx, cdj_mlp* mlp, const float* params) {
  if (!ctx || !mlp || !params) return CDJ_ERR_ARG;
  HIPCHK(hipMemcpy(mlp->m.params, params, sizeof(float) * mlp->m.nparams(), hipMemcpyHostToDevice));
  return CDJ_OK;
}

}  // extern "C"

namespace {

struct MeshGuard {
  MeshFeat f;
  ~MeshGuard() { mesh_free(f); }
};

// loss/gradient of one batch of samples on prepared mesh features; leaves Y (acts) and grads on the device
int mlp_step(cdj_ctx* ctx, cdj_mlp* mlp, const MeshFeat& f, const double* jac, const double* B, const double* X,
             int S, double invG, DBuf& acts, DBuf& dY, DBuf& work, DBuf& E, DBuf& lossp, double* loss,
             hipStream_t st) {
  const Mlp& m = mlp->m;
  const int64_t n = f.n;
  HIPCHK(mlp_forward(m, f.feat, n, acts.as<float>(), st));
  const float* Y = acts.as<float>() + 2 * n * m.H;
  HIPCHK(spai_loss_grad(jac, Y, f, invG, B, X, S, E.as<double>(), dY.as<float>(), lossp.as<double>(), st));
  HIPCHK(mlp_backward(m, f.feat, n, acts.as<float>(), dY.as<float>(), work.as<float>(), st));
  if (loss) {
    DBuf tot;
    HIPCHK(tot.alloc(sizeof(double)));
    HIPCHK(device_sum(lossp.as<double>(), n, tot.as<double>(), st));
    HIPCHK(hipMemcpyAsync(loss, tot.p, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return CDJ_OK;
}

}  // namespace

extern "C" {

int cdj_mlp_train_dev(cdj_ctx* ctx, cdj_mlp* mlp, const double* P1, const double* P2, const double* P3, int64_t n,
                      double G, const double* A, int64_t lda, int32_t nsamples, int32_t iters, float lr,
                      uint64_t seed, double* loss, void* stream) {
  if (!ctx || !mlp || !A) return CDJ_ERR_ARG;
  if (n <= 0 || lda < 3 * n || nsamples < 1 || iters < 1 || G <= 0) return fail(ctx, CDJ_ERR_ARG, "invalid training arguments");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  const Mlp& m = mlp->m;
  const int S = nsamples;
  const int64_t N3 = 3 * n;
  MeshGuard g;
  HIPCHK(mesh_prepare(P1, P2, P3, n, m.k, g.f, st));
  DBuf jac, X, B, part, acts, dY, work, E, lossp;
  HIPCHK(jac.alloc(sizeof(double) * 9 * n));
  HIPCHK(X.alloc(sizeof(double) * N3 * S));
  HIPCHK(B.alloc(sizeof(double) * N3 * S));
  HIPCHK(part.alloc(gemv_part_bytes(N3, N3)));
  HIPCHK(acts.alloc(sizeof(float) * n * (2 * m.H + m.O)));
  HIPCHK(dY.alloc(sizeof(float) * n * m.O));
  HIPCHK(work.alloc(sizeof(float) * 2 * n * m.H));
  HIPCHK(E.alloc(sizeof(double) * n * S * 3));
  HIPCHK(lossp.alloc(sizeof(double) * n));
  HIPCHK(jacobi_blocks(A, n, lda, jac.as<double>(), st));
  HIPCHK(randn(X.as<double>(), N3 * S, seed, st));
  for (int s = 0; s < S; ++s)
    gemv_splitk<double>(A, N3, N3, lda, X.as<double>() + s * N3, B.as<double>() + s * N3, 1.0, nullptr, 0.0,
                        part.as<double>(), st);
  for (int it = 0; it < iters; ++it) {
    int rc = mlp_step(ctx, mlp, g.f, jac.as<double>(), B.as<double>(), X.as<double>(), S, 1.0 / G, acts, dY, work, E,
                      lossp, it == iters - 1 ? loss : nullptr, st);
    if (rc) return rc;
    HIPCHK(mlp_adam(mlp->m, lr, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

int cdj_mlp_loss_grad(cdj_ctx* ctx, cdj_mlp* mlp, const double* P1, const double* P2, const double* P3, int64_t n,
                      double G, const double* jac, const double* B, const double* X, int32_t nsamples, double* loss,
                      float* grad, float* feat_out, int32_t* nbr_out, float* geo_out) {
  if (!ctx || !mlp || !P1 || !P2 || !P3 || !jac || !B || !X || !loss || !grad) return CDJ_ERR_ARG;
  if (n <= 0 || nsamples < 1 || G <= 0) return fail(ctx, CDJ_ERR_ARG, "invalid arguments");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const Mlp& m = mlp->m;
  const int S = nsamples;
  const int64_t N3 = 3 * n;
  DBuf d1, d2, d3, dj, dB, dX, acts, dY, work, E, lossp;
  HIPCHK(d1.alloc(sizeof(double) * 3 * n));
  HIPCHK(d2.alloc(sizeof(double) * 3 * n));
  HIPCHK(d3.alloc(sizeof(double) * 3 * n));
  HIPCHK(dj.alloc(sizeof(double) * 9 * n));
  HIPCHK(dB.alloc(sizeof(double) * N3 * S));
  HIPCHK(dX.alloc(sizeof(double) * N3 * S));
  HIPCHK(hipMemcpyAsync(d1.p, P1, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d2.p, P2, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d3.p, P3, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dj.p, jac, sizeof(double) * 9 * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dB.p, B, sizeof(double) * N3 * S, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dX.p, X, sizeof(double) * N3 * S, hipMemcpyHostToDevice, st));
  MeshGuard g;
  HIPCHK(mesh_prepare(d1.as<double>(), d2.as<double>(), d3.as<double>(), n, m.k, g.f, st));
  HIPCHK(acts.alloc(sizeof(float) * n * (2 * m.H + m.O)));
  HIPCHK(dY.alloc(sizeof(float) * n * m.O));
  HIPCHK(work.alloc(sizeof(float) * 2 * n * m.H));
  HIPCHK(E.alloc(sizeof(double) * n * S * 3));
  HIPCHK(lossp.alloc(sizeof(double) * n));
  int rc = mlp_step(ctx, mlp, g.f, dj.as<double>(), dB.as<double>(), dX.as<double>(), S, 1.0 / G, acts, dY, work, E,
                    lossp, loss, st);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(grad, m.grad, sizeof(float) * m.nparams(), hipMemcpyDeviceToHost, st));
  if (feat_out) HIPCHK(hipMemcpyAsync(feat_out, g.f.feat, sizeof(float) * n * m.F, hipMemcpyDeviceToHost, st));
  if (nbr_out) HIPCHK(hipMemcpyAsync(nbr_out, g.f.nbr, sizeof(int32_t) * n * m.k, hipMemcpyDeviceToHost, st));
  if (geo_out) HIPCHK(hipMemcpyAsync(geo_out, g.f.geo, sizeof(float) * n * GEO, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

int cdj_precond_solve_dev(cdj_ctx* ctx, const cdj_mlp* mlp, int32_t pc, const double* P1, const double* P2,
                          const double* P3, int64_t n, double G, const double* A, int64_t lda, const double* b,
                          double* x, const cdj_opts* opts, cdj_report* report, void* stream) {
  if (!ctx || !A || !b || !x) return CDJ_ERR_ARG;
  const bool learned = pc == CDJ_PC_LEARNED || pc == CDJ_PC_LEARNED_BF16;
  if (learned && !mlp) return fail(ctx, CDJ_ERR_ARG, "learned preconditioner needs a model");
  if (n <= 0 || lda < 3 * n) return fail(ctx, CDJ_ERR_DIMS, "invalid matrix dimensions");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  const int64_t N3 = 3 * n;
  if (pc == CDJ_PC_LU32) {
    cdj_opts o = opts ? *opts : cdj_opts{CDJ_PREC_MIXED32, 0, 1e-10, 0, CDJ_ASM_FP64, 2.0};
    return solve_mixed32(ctx, A, N3, lda, b, x, 1, &o, report, st);
  }
  GmresOpts go;
  go.restart = 60;
  go.max_iter = 3000;
  go.max_restarts = 200;
  if (opts && opts->tol > 0) go.tol = opts->tol;
  if (opts && opts->max_iter > 0) go.max_iter = opts->max_iter;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  MeshGuard g;
  DBuf jac, acts, wbf;
  const float* Y = nullptr;
  if (pc == CDJ_PC_JACOBI || learned) {
    HIPCHK(jac.alloc(sizeof(double) * 9 * n));
    HIPCHK(jacobi_blocks(A, n, lda, jac.as<double>(), st));
  }
  if (learned) {
    const Mlp& m = mlp->m;
    HIPCHK(mesh_prepare(P1, P2, P3, n, m.k, g.f, st));
    HIPCHK(acts.alloc(sizeof(float) * n * (2 * m.H + m.O)));
    if (pc == CDJ_PC_LEARNED_BF16) {
      HIPCHK(wbf.alloc(sizeof(unsigned short) * mlp_bf16_weight_count(m)));
      HIPCHK(mlp_forward_bf16(m, g.f.feat, n, acts.as<float>(), wbf.as<unsigned short>(), st));
    } else {
      HIPCHK(mlp_forward(m, g.f.feat, n, acts.as<float>(), st));
    }
    Y = acts.as<float>() + 2 * n * m.H;
  }
  const double invG = 1.0 / G;
  Precond M = [&](const double* r, double* z) -> hipError_t {
    if (pc == CDJ_PC_NONE) return hipMemcpyAsync(z, r, sizeof(double) * N3, hipMemcpyDeviceToDevice, st);
    if (pc == CDJ_PC_JACOBI) {
      MeshFeat f0;
      f0.n = n;
      f0.k = 0;
      return spai_apply(jac.as<double>(), nullptr, f0, invG, r, z, st);
    }
    return spai_apply(jac.as<double>(), Y, g.f, invG, r, z, st);
  };
  // CDJ_PREC_MIXED32: Arnoldi products on an FP32 copy of A, FP64 residual at every restart
  DBuf A32;
  if (opts && opts->precision == CDJ_PREC_MIXED32) {
    HIPCHK(A32.alloc(sizeof(float) * N3 * N3));
    convert_matrix_d2f(A, lda, A32.as<float>(), N3, N3, N3, st);
    go.A32 = A32.as<float>();
    go.lda32 = N3;
  }
  GmresResult gr;
  HIPCHK(fgmres(A, N3, lda, M, b, x, go, gr, st));
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (report) {
    report->path = CDJ_PREC_LEARNED;
    report->iterations = gr.iterations;
    report->residual = gr.residual;
    report->solve_ms = ms;
  }
  if (!gr.converged) {
    // fallback: FP32-LU preconditioned refinement (itself falling back to the FP64 LU)
    cdj_opts o = opts ? *opts : cdj_opts{CDJ_PREC_MIXED32, 0, 1e-10, 0, CDJ_ASM_FP64, 2.0};
    o.max_iter = 0;
    cdj_report r2{};
    int rc = solve_mixed32(ctx, A, N3, lda, b, x, 1, &o, &r2, st);
    if (rc) return rc;
    if (report) {
      report->path = r2.path;
      report->iterations += r2.iterations;
      report->residual = r2.residual;
      report->solve_ms += r2.solve_ms;
    }
  }
  return CDJ_OK;
}

}  // extern "C"

// ---- batched assembly and solve ----------------------------------------------------------------------------
//
// A plan holds the device buffers of one batch shape (triangle counts, right-hand-side counts, half-space flag,
// assembly mode, elastic constants), the tile shape chosen by the autotuner and the captured graph of the
// assembly sequence.  A call with the same shape reuses all three; another shape replaces the plan.

namespace {

struct TuneKey {
  int hs, mixed, bucket;
  bool operator<(const TuneKey& o) const {
    return hs != o.hs ? hs < o.hs : (mixed != o.mixed ? mixed < o.mixed : bucket < o.bucket);
  }
};
std::map<TuneKey, std::pair<int, int>>& tune_cache() {
  static std::map<TuneKey, std::pair<int, int>> m;
  return m;
}

}  // namespace

struct BatchPlan {
  std::vector<int64_t> n, nrhs;
  int hs = 0;
  bool mixed = false;
  double near_factor = 2.0, nu = 0, lam = 0, G = 0;
  int64_t ntot = 0, a_total = 0, b_total = 0;
  std::vector<BatchMeshDev> mh;      // host copy of the mesh table
  std::vector<int64_t> a_off, b_off; // per mesh: A (3n x (3n + nrhs)) and B/X (3n x nrhs) offsets, in doubles
  DBuf geo, recv, rec, rec32, near, meshes, items, counters, A, B, X;
  int tpb = 0, wgs = 0;
  int64_t nitems = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // solve phase: precision (CDJ_PREC_*), buffers for this batch shape, and for the mixed precisions the captured
  // graph of the solve's fixed part (FP32 LUs, permutations, first solve and one correction of every mesh)
  int solve_prec = CDJ_PREC_FP64;
  bool solve_layout = false, bufs_mixed = false, bufs_fp64 = false, device_perm = false;
  std::vector<int64_t> a32_off, v_off, q_off;  // per mesh: FP32 copy of A, pivots / permutation, column norms
  std::vector<size_t> ws_off;                  // per mesh: LU workspace
  int64_t a32_total = 0, v_total = 0, q_total = 0, n3_max = 0, napply = 0;
  size_t ws_total = 0, part_bytes = 0;
  // nrm: ||R||^2 per column (latest residual), ||B||^2 per column, ||R||^2 after the first solve
  // jobs0 / jobs1: factor applications to B (X = M^-1 B) and to R (X += M^-1 R); jobsx: a round's subset
  DBuf A32, ipiv, perm, info, W, R, nrm, lu_ws, part, jobs0, jobs1, jobsx, pjobs;
  std::vector<LuApplyJob> j1_host;  // host copy of jobs1
  std::vector<int64_t> jfirst;      // mesh m's applications are jobs [jfirst[m], jfirst[m + 1])
  hipGraph_t sgraph = nullptr;
  hipGraphExec_t sexec = nullptr;
  ~BatchPlan() {
    if (sexec) (void)hipGraphExecDestroy(sexec);
    if (sgraph) (void)hipGraphDestroy(sgraph);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  bool same(const std::vector<int64_t>& n2, const std::vector<int64_t>& r2, int hs2, bool mixed2, double f2,
            double nu2, double lam2, double G2, int prec2) const {
    return n == n2 && nrhs == r2 && hs == hs2 && mixed == mixed2 && near_factor == f2 && nu == nu2 && lam == lam2 &&
           G == G2 && solve_prec == prec2;
  }
};

void batch_plan_free(BatchPlan* p) { delete p; }

namespace {

// item prefix sums for a source-chunk width tpb
std::vector<int64_t> batch_items(const std::vector<int64_t>& n, int tpb) {
  std::vector<int64_t> st(n.size() + 1, 0);
  for (size_t m = 0; m < n.size(); ++m) st[m + 1] = st[m] + ((n[m] + 255) / 256) * ((n[m] + tpb - 1) / tpb);
  return st;
}

// The assembly sequence of a plan: counters cleared, receivers and records built, far (and near) passes.
hipError_t batch_enqueue(BatchPlan& p, hipStream_t st) {
  const int64_t N = p.ntot;
  const double* P1 = p.geo.as<double>();
  const double* P2 = P1 + 3 * N;
  const double* P3 = P2 + 3 * N;
  const double* nrm = P3 + 3 * N;
  const double* mid = nrm + 3 * N;
  unsigned long long* cnt = p.counters.as<unsigned long long>();
  launch_batch_clear(cnt, 4, st);
  launch_receivers(nrm, mid, N, p.recv.as<double>(), st);
  launch_setup(P1, P2, P3, N, p.hs, p.rec.as<double>(), st);
  const int nmesh = (int)p.n.size();
  if (p.mixed) {
    const int RS = p.hs ? REC_HS : REC_FS;
    launch_rec_to_f32(p.rec.as<double>(), p.rec32.as<float>(), (int64_t)RS * N, st);
    launch_near_info(P1, P2, P3, N, p.near_factor, p.near.as<double>(), st);
    launch_batch_far(p.recv.as<double>(), p.rec32.p, true, p.hs, p.meshes.as<BatchMeshDev>(), p.items.as<int64_t>(),
                     nmesh, p.tpb, p.wgs, p.nu, p.lam, p.G, p.A.as<double>(), cnt, st);
    launch_batch_near(p.recv.as<double>(), p.rec.as<double>(), p.near.as<double>(), p.hs, p.meshes.as<BatchMeshDev>(),
                      p.items.as<int64_t>(), nmesh, p.tpb, p.wgs, p.nu, p.lam, p.G, p.A.as<double>(), cnt + 1,
                      cnt + 2, st);
  } else {
    launch_batch_far(p.recv.as<double>(), p.rec.p, false, p.hs, p.meshes.as<BatchMeshDev>(), p.items.as<int64_t>(),
                     nmesh, p.tpb, p.wgs, p.nu, p.lam, p.G, p.A.as<double>(), cnt, st);
  }
  return hipGetLastError();
}

// Autotuner: times the assembly for each (source chunk, persistent workgroups) candidate on this batch and keeps
// the fastest; results are cached per (half-space, mode, size bucket) for later plans.
int batch_tune(cdj_ctx* ctx, BatchPlan& p, hipStream_t st) {
  int64_t pairs = 0;
  for (int64_t v : p.n) pairs += v * v;
  int bucket = 0;
  while ((1ll << bucket) < pairs) ++bucket;
  const TuneKey key{p.hs, p.mixed ? 1 : 0, bucket};
  auto it = tune_cache().find(key);
  int cus = 256;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
      cus = prop.multiProcessorCount;
  }
  std::vector<std::pair<int, int>> cands;
  if (it != tune_cache().end()) {
    cands.push_back(it->second);
  } else {
    const int tpbs_fs[] = {8, 16, 32, 64}, tpbs_hs[] = {2, 4, 8, 16};
    for (int t : (p.hs ? tpbs_hs : tpbs_fs))
      for (int w : {cus, 2 * cus, 4 * cus}) cands.push_back({t, w});
  }
  float best = 1e30f;
  std::pair<int, int> choice = cands[0];
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (auto c : cands) {
    p.tpb = c.first;
    p.wgs = c.second;
    const std::vector<int64_t> items = batch_items(p.n, p.tpb);
    HIPCHK(hipMemcpyAsync(p.items.p, items.data(), sizeof(int64_t) * items.size(), hipMemcpyHostToDevice, st));
    if (cands.size() == 1) break;
    HIPCHK(batch_enqueue(p, st));  // warm-up (first touch of the code object)
    HIPCHK(hipEventRecord(e0, st));
    HIPCHK(batch_enqueue(p, st));
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      choice = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p.tpb = choice.first;
  p.wgs = choice.second;
  const std::vector<int64_t> items = batch_items(p.n, p.tpb);
  p.nitems = items.back();
  HIPCHK(hipMemcpyAsync(p.items.p, items.data(), sizeof(int64_t) * items.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  tune_cache()[key] = choice;
  return CDJ_OK;
}

int batch_plan_build(cdj_ctx* ctx, const cdj_mesh* ms, int64_t nmesh, int hs, bool mixed, double f, double nu,
                     double lam, double G, int prec, hipStream_t st) {
  std::vector<int64_t> n(nmesh), r(nmesh);
  for (int64_t m = 0; m < nmesh; ++m) {
    n[m] = ms[m].n;
    r[m] = ms[m].nrhs;
  }
  if (ctx->batch && ctx->batch->same(n, r, hs, mixed, f, nu, lam, G, prec)) return CDJ_OK;
  batch_plan_free(ctx->batch);
  ctx->batch = nullptr;
  // built aside and installed only once every allocation has succeeded, so a failed build never leaves a plan
  // that a later call of the same shape would take for complete
  std::unique_ptr<BatchPlan> plan(new BatchPlan());
  BatchPlan& p = *plan;
  p.n = n;
  p.nrhs = r;
  p.hs = hs;
  p.mixed = mixed;
  p.near_factor = f;
  p.solve_prec = prec;
  p.nu = nu;
  p.lam = lam;
  p.G = G;
  p.mh.resize(nmesh);
  p.a_off.resize(nmesh);
  p.b_off.resize(nmesh);
  for (int64_t m = 0; m < nmesh; ++m) {
    const int64_t N3 = 3 * n[m];
    p.mh[m] = BatchMeshDev{p.ntot, n[m], p.a_total};
    p.a_off[m] = p.a_total;
    p.b_off[m] = p.b_total;
    p.ntot += n[m];
    p.a_total += N3 * (N3 + r[m]);
    p.b_total += N3 * r[m];
  }
  const int RS = hs ? REC_HS : REC_FS;
  HIPCHK(p.geo.alloc(sizeof(double) * 15 * p.ntot));
  HIPCHK(p.recv.alloc(sizeof(double) * RECV * p.ntot));
  HIPCHK(p.rec.alloc(sizeof(double) * RS * p.ntot));
  if (mixed) {
    HIPCHK(p.rec32.alloc(sizeof(float) * RS * p.ntot));
    HIPCHK(p.near.alloc(sizeof(double) * NEAR_REC * p.ntot));
  }
  HIPCHK(p.meshes.alloc(sizeof(BatchMeshDev) * nmesh));
  HIPCHK(p.items.alloc(sizeof(int64_t) * (nmesh + 1)));
  HIPCHK(p.counters.alloc(sizeof(unsigned long long) * 4));
  HIPCHK(p.A.alloc(sizeof(double) * p.a_total));
  HIPCHK(p.B.alloc(sizeof(double) * p.b_total));
  HIPCHK(p.X.alloc(sizeof(double) * p.b_total));
  HIPCHK(hipMemcpyAsync(p.meshes.p, p.mh.data(), sizeof(BatchMeshDev) * nmesh, hipMemcpyHostToDevice, st));
  ctx->batch = plan.release();
  return CDJ_OK;
}

// ---- solve phase of a batch ---------------------------------------------------------------------------------
// The meshes are independent systems of uneven size.  Their LUs are stepped together by panel column, so that one
// cooperative launch factors the panels of many meshes (lu_factor_batched); the refinement applies every mesh's
// factors in one launch per round, with one host synchronisation per round for all meshes.
constexpr int BATCH_IR_ROUNDS = 10;  // refinement rounds before a mesh moves to FGMRES-IR

int batch_solve_setup(cdj_ctx* ctx, BatchPlan& p, bool mixed) {
  const int64_t nmesh = (int64_t)p.n.size();
  if (!p.solve_layout) {
    p.a32_off.assign(nmesh, 0);
    p.v_off.assign(nmesh, 0);
    p.q_off.assign(nmesh, 0);
    p.ws_off.assign(nmesh, 0);
    for (int64_t m = 0; m < nmesh; ++m) {
      const int64_t N3 = 3 * p.n[m];
      p.a32_off[m] = p.a32_total;
      p.v_off[m] = p.v_total;
      p.q_off[m] = p.q_total;
      p.ws_off[m] = p.ws_total;
      p.a32_total += N3 * N3;
      p.v_total += N3;
      p.q_total += p.nrhs[m];
      p.ws_total += (lu_workspace_bytes(N3) + 255) / 256 * 256;
    }
    p.solve_layout = true;
  }
  if (!p.bufs_fp64) {
    HIPCHK(p.lu_ws.alloc(p.ws_total));
    HIPCHK(p.ipiv.alloc(sizeof(int32_t) * p.v_total));
    HIPCHK(p.info.alloc(sizeof(int) * nmesh));
    p.bufs_fp64 = true;
  }
  if (mixed && !p.bufs_mixed) {
    size_t pb = 0;
    for (int64_t m = 0; m < nmesh; ++m) pb = std::max(pb, residual_multi_bytes(3 * p.n[m], p.nrhs[m]));
    p.part_bytes = pb;
    HIPCHK(p.part.alloc(p.part_bytes));
    HIPCHK(p.A32.alloc(sizeof(float) * p.a32_total));
    HIPCHK(p.perm.alloc(sizeof(int32_t) * p.v_total));
    HIPCHK(p.W.alloc(sizeof(float) * p.b_total));
    HIPCHK(p.R.alloc(sizeof(double) * p.b_total));
    HIPCHK(p.nrm.alloc(sizeof(double) * 3 * p.q_total));
    // factor-application and permutation job tables, fixed for the plan (uploaded once, outside any capture)
    std::vector<LuApplyJob> j0, j1;
    std::vector<PermJob> pj;
    for (int64_t m = 0; m < nmesh; ++m) {
      const int64_t N3 = 3 * p.n[m], nr = p.nrhs[m];
      float* LUm = p.A32.as<float>() + p.a32_off[m];
      int32_t* pm = p.perm.as<int32_t>() + p.v_off[m];
      pj.push_back(PermJob{p.ipiv.as<int32_t>() + p.v_off[m], pm, N3});
      for (int64_t q0 = 0; q0 < nr; q0 += APPLY_NR) {
        const int64_t o = p.b_off[m] + q0 * N3;
        const int32_t nq = (int32_t)std::min<int64_t>(APPLY_NR, nr - q0);
        j0.push_back(LuApplyJob{LUm, pm, p.B.as<double>() + o, p.X.as<double>() + o, p.W.as<float>() + o, N3, N3, nq, 0});
        j1.push_back(LuApplyJob{LUm, pm, p.R.as<double>() + o, p.X.as<double>() + o, p.W.as<float>() + o, N3, N3, nq, 1});
      }
    }
    p.napply = (int64_t)j0.size();
    p.j1_host = j1;
    p.jfirst.assign(nmesh + 1, 0);
    for (int64_t m = 0; m < nmesh; ++m) p.jfirst[m + 1] = p.jfirst[m] + (p.nrhs[m] + APPLY_NR - 1) / APPLY_NR;
    HIPCHK(p.jobs0.alloc(sizeof(LuApplyJob) * j0.size()));
    HIPCHK(p.jobs1.alloc(sizeof(LuApplyJob) * j1.size()));
    HIPCHK(p.jobsx.alloc(sizeof(LuApplyJob) * j1.size()));
    HIPCHK(p.pjobs.alloc(sizeof(PermJob) * pj.size()));
    HIPCHK(hipMemcpy(p.jobs0.p, j0.data(), sizeof(LuApplyJob) * j0.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p.jobs1.p, j1.data(), sizeof(LuApplyJob) * j1.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p.pjobs.p, pj.data(), sizeof(PermJob) * pj.size(), hipMemcpyHostToDevice));
    for (int64_t m = 0; m < nmesh; ++m) p.n3_max = std::max(p.n3_max, 3 * p.n[m]);
    // the one-time kernel set-ups happen here, never inside a capture
    if (lu_batched_capacity<float>() <= 0) return fail(ctx, CDJ_ERR_DEVICE, "batched LU set-up failed");
    p.device_perm = p.n3_max <= pivots_to_perm_max();
    p.bufs_mixed = true;
  }
  return CDJ_OK;
}

// Meshes in decreasing size (largest panels first in each launch).
std::vector<int64_t> batch_solve_order(const BatchPlan& p) {
  std::vector<int64_t> order(p.n.size());
  for (size_t m = 0; m < order.size(); ++m) order[m] = (int64_t)m;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return p.n[a] > p.n[b]; });
  return order;
}

int batch_lu_status(cdj_ctx* ctx, const std::vector<int>& info, bool singular_is_error) {
  for (size_t m = 0; m < info.size(); ++m) {
    if (info[m] < 0) return fail(ctx, CDJ_ERR_DEVICE, "LU panel barrier wait timed out");
    if (info[m] > 0 && singular_is_error)
      return fail(ctx, CDJ_ERR_SINGULAR, "matrix is exactly singular (mesh " + std::to_string(m) + ", zero pivot " +
                                             std::to_string(info[m]) + ")");
  }
  return CDJ_OK;
}

// FP64: [A | B] of every mesh factored in place (B carried along as extra columns), then U X = L^-1 P B.
int batch_solve_fp64(cdj_ctx* ctx, BatchPlan& p, const cdj_mesh* meshes, int64_t nmesh, cdj_report* reports,
                     double* solve_ms, hipStream_t st) {
  if (int rc = batch_solve_setup(ctx, p, false)) return rc;
  const std::vector<int64_t> order = batch_solve_order(p);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  std::vector<LuJob<double>> jobs;
  for (int64_t m : order) {
    const int64_t N3 = 3 * meshes[m].n, nr = meshes[m].nrhs;
    double* Am = p.A.as<double>() + p.a_off[m];
    HIPCHK(hipMemcpyAsync(Am + N3 * N3, p.B.as<double>() + p.b_off[m], sizeof(double) * N3 * nr,
                          hipMemcpyDeviceToDevice, st));
    jobs.push_back(LuJob<double>{Am, N3, N3 + nr, N3, p.ipiv.as<int32_t>() + p.v_off[m], p.info.as<int>() + m,
                                 p.lu_ws.as<char>() + p.ws_off[m]});
  }
  HIPCHK(lu_factor_batched<double>(jobs.data(), (int)jobs.size(), st));
  for (int64_t m : order) {
    const int64_t N3 = 3 * meshes[m].n, nr = meshes[m].nrhs;
    double* Am = p.A.as<double>() + p.a_off[m];
    HIPCHK(tri_solve_upper<double>(Am, N3, N3, Am + N3 * N3, nr, N3, st));
    HIPCHK(hipMemcpyAsync(p.X.as<double>() + p.b_off[m], Am + N3 * N3, sizeof(double) * N3 * nr,
                          hipMemcpyDeviceToDevice, st));
  }
  HIPCHK(hipEventRecord(e1, st));
  std::vector<int> info(nmesh);
  HIPCHK(hipMemcpyAsync(info.data(), p.info.p, sizeof(int) * nmesh, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (int rc = batch_lu_status(ctx, info, true)) return rc;
  *solve_ms = ms;
  if (reports)
    for (int64_t m = 0; m < nmesh; ++m) {
      reports[m].path = CDJ_PREC_FP64;
      reports[m].iterations = 0;
      reports[m].residual = -1.0;
      reports[m].solve_ms = ms;
    }
  return CDJ_OK;
}

// Mixed: FP32 copy of each A factored on MFMA, X = M^-1 B for all of a mesh's right-hand sides at once, then
// rounds of X += M^-1 (B - A X) with the FP64 residual, until every column reaches the tolerance.  A mesh whose
// residual stops falling (less than halved by a round) or whose FP32 factor has a zero pivot moves to FGMRES-IR
// per column, and from there to the FP64 LU, as in the single-mesh solve.
//
// The fixed part (FP32 copies and LUs of all meshes, the permutations, the first solve and one correction, each
// followed by its FP64 residual) is only kernel launches on st: with device_perm it is captured in the plan's solve
// graph.  Without it (a mesh above PERM_MAX unknowns) the permutations come from the host between the two halves.
hipError_t batch_solve_lu(BatchPlan& p, hipStream_t st) {
  const std::vector<int64_t> order = batch_solve_order(p);
  double* nrm = p.nrm.as<double>();
  std::vector<LuJob<float>> jobs;
  for (int64_t m : order) {
    const int64_t N3 = 3 * p.n[m];
    convert_matrix_d2f(p.A.as<double>() + p.a_off[m], N3, p.A32.as<float>() + p.a32_off[m], N3, N3, N3, st);
    colnorm2(p.B.as<double>() + p.b_off[m], N3, N3, p.nrhs[m], nrm + p.q_total + p.q_off[m], st);
    jobs.push_back(LuJob<float>{p.A32.as<float>() + p.a32_off[m], N3, N3, N3, p.ipiv.as<int32_t>() + p.v_off[m],
                                p.info.as<int>() + m, p.lu_ws.as<char>() + p.ws_off[m]});
  }
  return lu_factor_batched<float>(jobs.data(), (int)jobs.size(), st, p.solve_prec == CDJ_PREC_MIXED16);
}

hipError_t batch_residuals(BatchPlan& p, double* nrm_out, hipStream_t st) {
  hipError_t e = hipSuccess;
  for (size_t m = 0; e == hipSuccess && m < p.n.size(); ++m) {
    const int64_t N3 = 3 * p.n[m];
    const int64_t o = p.b_off[m];
    e = residual_multi(p.A.as<double>() + p.a_off[m], N3, N3, p.B.as<double>() + o, N3, p.X.as<double>() + o, N3,
                       p.R.as<double>() + o, N3, p.nrhs[m], p.part.as<double>(), nrm_out + p.q_off[m], st);
  }
  return e;
}

hipError_t batch_solve_apply(BatchPlan& p, hipStream_t st) {
  double* nrm = p.nrm.as<double>();
  hipError_t e = lu_apply_batched(p.jobs0.as<LuApplyJob>(), (int)p.napply, st);
  if (e == hipSuccess) e = batch_residuals(p, nrm + 2 * p.q_total, st);
  if (e == hipSuccess) e = lu_apply_batched(p.jobs1.as<LuApplyJob>(), (int)p.napply, st);
  if (e == hipSuccess) e = batch_residuals(p, nrm, st);
  return e;
}

// the whole fixed part, for the capture
hipError_t batch_solve_enqueue(BatchPlan& p, hipStream_t st) {
  hipError_t e = batch_solve_lu(p, st);
  if (e == hipSuccess)
    e = pivots_to_perm(p.pjobs.as<PermJob>(), (int)p.n.size(), p.n3_max, st);
  if (e == hipSuccess) e = batch_solve_apply(p, st);
  return e;
}

// The permutations on the host (for meshes too large for the device kernel): synchronises the stream.
int batch_host_perms(cdj_ctx* ctx, BatchPlan& p, hipStream_t st) {
  std::vector<int32_t> hp(p.v_total);
  HIPCHK(hipMemcpyAsync(hp.data(), p.ipiv.p, sizeof(int32_t) * p.v_total, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (size_t m = 0; m < p.n.size(); ++m) {
    const int64_t N3 = 3 * p.n[m];
    int32_t* v = hp.data() + p.v_off[m];
    std::vector<int32_t> q(N3);
    for (int64_t i = 0; i < N3; ++i) q[i] = (int32_t)i;
    for (int64_t i = 0; i < N3; ++i) std::swap(q[i], q[v[i]]);
    std::copy(q.begin(), q.end(), v);
  }
  HIPCHK(hipMemcpyAsync(p.perm.p, hp.data(), sizeof(int32_t) * p.v_total, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

// After the fixed part: status, refinement rounds while needed, fallbacks, reports.
int batch_solve_mixed_finish(cdj_ctx* ctx, BatchPlan& p, const cdj_mesh* meshes, int64_t nmesh,
                             const cdj_opts* opts, cdj_report* reports, double* worst, int* iters, int* path,
                             hipStream_t st) {
  GmresOpts go;
  if (opts && opts->tol > 0) go.tol = opts->tol;
  if (opts && opts->max_iter > 0) go.max_iter = opts->max_iter;
  const bool fp16 = p.solve_prec == CDJ_PREC_MIXED16;
  const std::vector<int64_t> order = batch_solve_order(p);
  double* nrm = p.nrm.as<double>();
  std::vector<int> info(nmesh);
  std::vector<double> hn(3 * p.q_total);
  HIPCHK(hipMemcpyAsync(info.data(), p.info.p, sizeof(int) * nmesh, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * 3 * p.q_total, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (int rc = batch_lu_status(ctx, info, false)) return rc;
  // worst relative residual of a mesh's columns from a norm section (NaN propagates)
  auto rel_of = [&](int64_t m, int64_t section) {
    double rel = 0.0;
    for (int64_t q = 0; q < meshes[m].nrhs; ++q) {
      const double rr = hn[section * p.q_total + p.q_off[m] + q], bb = hn[p.q_total + p.q_off[m] + q];
      const double r = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
      rel = std::isnan(r) || std::isnan(rel) ? NAN : std::max(rel, r);
    }
    return rel;
  };
  // 0: refining, 1: converged, 2: to the fallback.  Every mesh has had its first solve and one correction.
  std::vector<char> state(nmesh, 0);
  std::vector<int> rounds(nmesh, 1);
  std::vector<double> res(nmesh, 0.0), prev(nmesh, HUGE_VAL);
  for (int64_t m = 0; m < nmesh; ++m) {
    if (info[m] != 0) {
      state[m] = 2;
      rounds[m] = 0;
    } else {
      prev[m] = rel_of(m, 2);
    }
  }
  std::vector<LuApplyJob> hj;
  for (int round = 1;; ++round) {
    hj.clear();
    for (int64_t m : order) {
      if (state[m]) continue;
      const double rel = rel_of(m, 0);
      res[m] = rel;
      if (rel <= go.tol) {
        state[m] = 1;
        continue;
      }
      if (round >= BATCH_IR_ROUNDS || !(rel < 0.5 * prev[m])) {
        state[m] = 2;
        continue;
      }
      prev[m] = rel;
      rounds[m]++;
      for (int64_t t = p.jfirst[m]; t < p.jfirst[m + 1]; ++t) hj.push_back(p.j1_host[t]);
    }
    if (hj.empty()) break;
    HIPCHK(hipMemcpyAsync(p.jobsx.p, hj.data(), sizeof(LuApplyJob) * hj.size(), hipMemcpyHostToDevice, st));
    HIPCHK(lu_apply_batched(p.jobsx.as<LuApplyJob>(), (int)hj.size(), st));
    for (int64_t m : order) {
      if (state[m]) continue;
      const int64_t N3 = 3 * meshes[m].n, o = p.b_off[m];
      HIPCHK(residual_multi(p.A.as<double>() + p.a_off[m], N3, N3, p.B.as<double>() + o, N3, p.X.as<double>() + o, N3,
                            p.R.as<double>() + o, N3, meshes[m].nrhs, p.part.as<double>(), nrm + p.q_off[m], st));
    }
    HIPCHK(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * p.q_total, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  // fallbacks, one mesh at a time
  std::vector<int> mpath(nmesh, fp16 ? CDJ_PREC_MIXED16 : CDJ_PREC_MIXED32);
  std::vector<int> fit(nmesh, 0);
  for (int64_t m = 0; m < nmesh; ++m) {
    if (state[m] != 2) continue;
    const int64_t N3 = 3 * meshes[m].n, nr = meshes[m].nrhs;
    double* Am = p.A.as<double>() + p.a_off[m];
    double* Bm = p.B.as<double>() + p.b_off[m];
    double* Xm = p.X.as<double>() + p.b_off[m];
    bool ok = info[m] == 0;
    double w = 0.0;
    for (int64_t q = 0; ok && q < nr; ++q) {
      GmresResult gr;
      HIPCHK(fgmres_ir<float>(Am, N3, N3, p.A32.as<float>() + p.a32_off[m], N3, p.perm.as<int32_t>() + p.v_off[m],
                              Bm + q * N3, Xm + q * N3, go, gr, st));
      ok = gr.converged && !gr.trsv_timeout;
      w = std::max(w, gr.residual);
      fit[m] += gr.iterations;
    }
    if (ok) {
      res[m] = w;
      continue;
    }
    DBuf AB;
    HIPCHK(AB.alloc(sizeof(double) * N3 * (N3 + nr)));
    HIPCHK(hipMemcpyAsync(AB.p, Am, sizeof(double) * N3 * N3, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(AB.as<double>() + N3 * N3, Bm, sizeof(double) * N3 * nr, hipMemcpyDeviceToDevice, st));
    if (int rc = solve_augmented_fp64(ctx, AB.as<double>(), N3, N3, nr, nullptr, nullptr, nullptr, st)) return rc;
    HIPCHK(hipMemcpyAsync(Xm, AB.as<double>() + N3 * N3, sizeof(double) * N3 * nr, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    mpath[m] = CDJ_PREC_FP64;
    res[m] = -1.0;
  }
  for (int64_t m = 0; m < nmesh; ++m) {
    *worst = std::max(*worst, res[m]);
    *iters += rounds[m] + fit[m];
    // the batch's path is FP64 as soon as one mesh fell back (FP64 is the smallest code)
    *path = m == 0 ? mpath[m] : std::min(*path, mpath[m]);
    if (reports) {
      reports[m].path = mpath[m];
      reports[m].iterations = rounds[m] + fit[m];
      reports[m].residual = res[m];
    }
  }
  return CDJ_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

int batched_slip_run(cdj_ctx* ctx, const cdj_mesh* meshes, int64_t nmesh, double nu, double G, double lambda,
                     int32_t hs_flag, const cdj_opts* opts, cdj_report* reports, cdj_report* total) {
  if (nmesh < 0 || (nmesh > 0 && !meshes)) return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  bool mixed = false;
  double near_factor = 2.0;
  if (int rc = assembly_mode(ctx, opts, &mixed, &near_factor)) return rc;
  for (int64_t m = 0; m < nmesh; ++m) {
    const cdj_mesh& M = meshes[m];
    if (M.n <= 0 || M.nrhs < 1) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
    if (!M.P1 || !M.P2 || !M.P3 || !M.normals || !M.midpoints || !M.b || !M.x)
      return fail(ctx, CDJ_ERR_ARG, "NULL argument");
    if (hs_flag == 1)
      for (int64_t i = 0; i < M.n; ++i)
        if (M.P1[2 * M.n + i] > 0 || M.P2[2 * M.n + i] > 0 || M.P3[2 * M.n + i] > 0 || M.midpoints[2 * M.n + i] > 0)
          return fail(ctx, CDJ_ERR_HALFSPACE, "Half-space solution: Z coordinates must be negative!");
  }
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (nmesh == 0) return CDJ_OK;
  const int hs = hs_flag == 1;
  const int prec = opts && (opts->precision == CDJ_PREC_MIXED32 || opts->precision == CDJ_PREC_MIXED16)
                       ? opts->precision
                       : CDJ_PREC_FP64;
  if (int rc = batch_plan_build(ctx, meshes, nmesh, hs, mixed, near_factor, nu, lambda, G, prec, st)) return rc;
  BatchPlan& p = *ctx->batch;
  // geometry: five N x 3 column-major arrays over the concatenated triangles; the right-hand sides
  {
    std::vector<double> h(15 * (size_t)p.ntot);
    for (int64_t m = 0; m < nmesh; ++m) {
      const cdj_mesh& M = meshes[m];
      const double* src[5] = {M.P1, M.P2, M.P3, M.normals, M.midpoints};
      for (int a = 0; a < 5; ++a)
        for (int c = 0; c < 3; ++c)
          std::memcpy(&h[(size_t)a * 3 * p.ntot + (size_t)c * p.ntot + p.mh[m].tri0], src[a] + c * M.n,
                      sizeof(double) * M.n);
    }
    HIPCHK(hipMemcpyAsync(p.geo.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, st));
    for (int64_t m = 0; m < nmesh; ++m)
      HIPCHK(hipMemcpyAsync(p.B.as<double>() + p.b_off[m], meshes[m].b,
                            sizeof(double) * 3 * meshes[m].n * meshes[m].nrhs, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  // graphs: the assembly (after tuning); for the mixed precisions, the solve's fixed part as a second graph
  auto capture = [&](auto&& enqueue, hipGraph_t* g, hipGraphExec_t* x, const char* what) -> int {
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t ce = enqueue();
    hipError_t ee = hipStreamEndCapture(st, g);
    if (ce != hipSuccess) return hip_fail(ctx, ce, what);
    if (ee != hipSuccess) return hip_fail(ctx, ee, "hipStreamEndCapture");
    HIPCHK(hipGraphInstantiate(x, *g, nullptr, nullptr, 0));
    return CDJ_OK;
  };
  if (!p.exec) {
    if (int rc = batch_tune(ctx, p, st)) return rc;
    if (int rc = capture([&] { return batch_enqueue(p, st); }, &p.graph, &p.exec, "batch capture")) return rc;
  }
  const bool mixed_solve = prec != CDJ_PREC_FP64;
  if (mixed_solve) {
    if (int rc = batch_solve_setup(ctx, p, true)) return rc;
    if (p.device_perm && !p.sexec)
      if (int rc = capture([&] { return batch_solve_enqueue(p, st); }, &p.sgraph, &p.sexec, "batch solve capture"))
        return rc;
  }
  hipEvent_t e0, e1, e2;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventCreate(&e2));
  HIPCHK(hipEventRecord(e0, st));
  HIPCHK(hipGraphLaunch(p.exec, st));
  HIPCHK(hipEventRecord(e1, st));
  unsigned long long cnt[4] = {0, 0, 0, 0};
  double solve_ms = 0.0, worst = 0.0;
  int iters = 0, path = CDJ_PREC_FP64;
  int rc = CDJ_OK;
  if (mixed_solve) {
    if (p.sexec) {
      HIPCHK(hipGraphLaunch(p.sexec, st));
    } else {
      HIPCHK(batch_solve_lu(p, st));
      rc = batch_host_perms(ctx, p, st);
      if (!rc) HIPCHK(batch_solve_apply(p, st));
    }
    if (!rc) rc = batch_solve_mixed_finish(ctx, p, meshes, nmesh, opts, reports, &worst, &iters, &path, st);
    if (!rc) HIPCHK(hipEventRecord(e2, st));
  }
  HIPCHK(hipMemcpyAsync(cnt, p.counters.p, sizeof(cnt), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float ams = 0.f, sms = 0.f;
  (void)hipEventElapsedTime(&ams, e0, e1);
  if (mixed_solve && !rc) (void)hipEventElapsedTime(&sms, e1, e2);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  if (rc) return rc;
  ctx->near_entries = (int64_t)cnt[2];
  if (mixed_solve) {
    solve_ms = sms;
    if (reports)
      for (int64_t m = 0; m < nmesh; ++m) reports[m].solve_ms = solve_ms;
  } else {
    if (int rc2 = batch_solve_fp64(ctx, p, meshes, nmesh, reports, &solve_ms, st)) return rc2;
  }
  for (int64_t m = 0; m < nmesh; ++m) {
    HIPCHK(hipMemcpyAsync(meshes[m].x, p.X.as<double>() + p.b_off[m], sizeof(double) * 3 * meshes[m].n * meshes[m].nrhs,
                          hipMemcpyDeviceToHost, st));
    if (reports) {
      reports[m].assembly_ms = m == 0 ? ams : 0.0;
      reports[m].near_entries = 0;
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  if (total) {
    total->path = path;
    total->iterations = iters;
    total->residual = worst;
    total->assembly_ms = ams;
    total->solve_ms = solve_ms;
    total->near_entries = (int64_t)cnt[2];
  }
  return CDJ_OK;
}

}  // namespace

extern "C" {

int cdj_batched_slip(cdj_ctx* ctx, const cdj_mesh* meshes, int64_t nmesh, double nu, double G, double lambda,
                     int32_t hs_flag, const cdj_opts* opts, cdj_report* reports, cdj_report* total) {
  if (!ctx) return CDJ_ERR_ARG;
  const int rc = batched_slip_run(ctx, meshes, nmesh, nu, G, lambda, hs_flag, opts, reports, total);
  if (rc != CDJ_OK) {
    // whatever failed (an allocation, a capture, a solve), the next call starts from a fresh plan
    batch_plan_free(ctx->batch);
    ctx->batch = nullptr;
  }
  return rc;
}

int cdj_batch_tuning(const cdj_ctx* ctx, int32_t* tpb, int32_t* wgs) {
  if (!ctx || !ctx->batch) return CDJ_ERR_ARG;
  if (tpb) *tpb = ctx->batch->tpb;
  if (wgs) *wgs = ctx->batch->wgs;
  return CDJ_OK;
}

}  // extern "C"

// ---- traction rows at given receivers, least squares ----------------------------------------------------------
extern "C" {

int cdj_traction_rows(cdj_ctx* ctx, const double* points, const double* normals, int64_t m, const double* P1,
                      const double* P2, const double* P3, int64_t n, double nu, double G, double lambda,
                      int32_t hs_flag, double* const* T, const cdj_opts* opts) {
  if (!ctx) return CDJ_ERR_ARG;
  if (m < 0 || n < 0) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  if (!points || !normals || !P1 || !P2 || !P3) return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  for (int k = 0; k < 9; ++k)
    if (!T || !T[k]) return fail(ctx, CDJ_ERR_ARG, "traction output pointer is NULL");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (m == 0 || n == 0) return CDJ_OK;
  const int hs = hs_flag == 1;
  if (hs) {
    for (int64_t i = 0; i < m; ++i)
      if (points[2 * m + i] > 0) return fail(ctx, CDJ_ERR_HALFSPACE, "Half-space solution: Z coordinates must be negative!");
    for (int64_t i = 0; i < n; ++i)
      if (P1[2 * n + i] > 0 || P2[2 * n + i] > 0 || P3[2 * n + i] > 0)
        return fail(ctx, CDJ_ERR_HALFSPACE, "Half-space solution: Z coordinates must be negative!");
  }
  const int RS = hs ? REC_HS : REC_FS;
  DBuf dp, dn, d1, d2, d3, recv, rec, dT;
  HIPCHK(dp.alloc(sizeof(double) * 3 * m));
  HIPCHK(dn.alloc(sizeof(double) * 3 * m));
  HIPCHK(d1.alloc(sizeof(double) * 3 * n));
  HIPCHK(d2.alloc(sizeof(double) * 3 * n));
  HIPCHK(d3.alloc(sizeof(double) * 3 * n));
  HIPCHK(hipMemcpyAsync(dp.p, points, sizeof(double) * 3 * m, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dn.p, normals, sizeof(double) * 3 * m, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d1.p, P1, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d2.p, P2, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d3.p, P3, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  HIPCHK(recv.alloc(sizeof(double) * RECV * m));
  HIPCHK(rec.alloc(sizeof(double) * RS * n));
  const size_t vm = sizeof(double) * m * n;
  HIPCHK(dT.alloc(9 * vm));
  launch_receivers_at(dn.as<double>(), dp.as<double>(), m, recv.as<double>(), st);
  launch_setup(d1.as<double>(), d2.as<double>(), d3.as<double>(), n, hs, rec.as<double>(), st);
  OutPtrs<double> out{};
  out.ld = m;
  out.sign = 1.0;
  for (int k = 0; k < 9; ++k) out.p[k] = dT.as<double>() + (int64_t)k * m * n;
  const int tpb = (opts && opts->tile > 0) ? opts->tile : default_tile(hs, m);
  if (hs) {
    if (int rc = traction_hs_fp64(ctx, recv.as<double>(), m, d1.as<double>(), d2.as<double>(), d3.as<double>(),
                                  rec.as<double>(), n, tpb, nu, lambda, G, out, false, st))
      return rc;
  } else {
    launch_traction<double>(recv.as<double>(), m, rec.as<double>(), n, hs, tpb, nu, lambda, G, out, st);
  }
  HIPCHK(hipGetLastError());
  for (int k = 0; k < 9; ++k) HIPCHK(hipMemcpyAsync(T[k], out.p[k], vm, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

int cdj_lstsq(cdj_ctx* ctx, const double* A, int64_t m, int64_t k, const double* b, int64_t nrhs, double* x,
              cdj_report* report) {
  if (!ctx) return CDJ_ERR_ARG;
  if (m < 0 || k < 0 || nrhs < 0 || k > m) return fail(ctx, CDJ_ERR_DIMS, "invalid matrix dimensions");
  if (m + k > 65535) return fail(ctx, CDJ_ERR_DIMS, "least-squares system too large (m + k > 65535)");
  if (!A || !b || !x) return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (k == 0 || nrhs == 0) return CDJ_OK;
  DBuf dA, dB, dX;
  HIPCHK(dA.alloc(sizeof(double) * m * k));
  HIPCHK(dB.alloc(sizeof(double) * m * nrhs));
  HIPCHK(dX.alloc(sizeof(double) * k * nrhs));
  HIPCHK(hipMemcpyAsync(dA.p, A, sizeof(double) * m * k, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dB.p, b, sizeof(double) * m * nrhs, hipMemcpyHostToDevice, st));
  int rc = lstsq_dev(ctx, dA.as<double>(), m, k, m, dB.as<double>(), nrhs, dX.as<double>(), report, st);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(x, dX.p, sizeof(double) * k * nrhs, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

}  // extern "C"

// ---- crack-tip edges and stress intensities ---------------------------------------------------------------
namespace {

constexpr int64_t CT_MAX_TRI = 700000000;  // 3n edge indices fit an int

int crack_tip_check(cdj_ctx* ctx, int64_t n) {
  if (n < 0 || n > CT_MAX_TRI) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  return CDJ_OK;
}

}  // namespace

extern "C" {

int cdj_crack_tip_elements(cdj_ctx* ctx, const double* midpoints, const double* P1, const double* P2,
                           const double* P3, const double* normals, int64_t n, uint8_t* free_flags,
                           double* geometry) {
  if (!ctx) return CDJ_ERR_ARG;
  if (int rc = crack_tip_check(ctx, n)) return rc;
  if (!midpoints || !P1 || !P2 || !P3 || !normals || !free_flags || !geometry)
    return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (n == 0) return CDJ_OK;
  const size_t v3 = sizeof(double) * 3 * n;
  DBuf dm, d1, d2, d3, dn, slots, matched, flags, geom;
  HIPCHK(dm.alloc(v3));
  HIPCHK(d1.alloc(v3));
  HIPCHK(d2.alloc(v3));
  HIPCHK(d3.alloc(v3));
  HIPCHK(dn.alloc(v3));
  HIPCHK(slots.alloc(sizeof(int) * crack_tip_slots(n)));
  HIPCHK(matched.alloc(sizeof(int) * 3 * n));
  HIPCHK(flags.alloc(3 * n));
  HIPCHK(geom.alloc(sizeof(double) * 3 * n * CT_GEOM));
  HIPCHK(hipMemcpyAsync(dm.p, midpoints, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d1.p, P1, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d2.p, P2, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d3.p, P3, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dn.p, normals, v3, hipMemcpyHostToDevice, st));
  HIPCHK(crack_tip_elements(dm.as<double>(), d1.as<double>(), d2.as<double>(), d3.as<double>(), dn.as<double>(), n,
                            slots.as<int>(), matched.as<int>(), flags.as<uint8_t>(), geom.as<double>(), st));
  HIPCHK(hipMemcpyAsync(free_flags, flags.p, 3 * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(geometry, geom.p, sizeof(double) * 3 * n * CT_GEOM, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

int cdj_stress_intensity(cdj_ctx* ctx, const double* Dn, const double* Dss, const double* Dds, int64_t n, double G,
                         double nu, const double* normals, const uint8_t* free_flags, const double* geometry,
                         double* K) {
  if (!ctx) return CDJ_ERR_ARG;
  if (int rc = crack_tip_check(ctx, n)) return rc;
  if (!Dn || !Dss || !Dds || !normals || !free_flags || !geometry || !K) return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (n == 0) return CDJ_OK;
  DBuf dD, dn, flags, geom, dK;
  HIPCHK(dD.alloc(sizeof(double) * 3 * n));
  HIPCHK(dn.alloc(sizeof(double) * 3 * n));
  HIPCHK(flags.alloc(3 * n));
  HIPCHK(geom.alloc(sizeof(double) * 3 * n * CT_GEOM));
  HIPCHK(dK.alloc(sizeof(double) * 3 * n * CT_K));
  double* D = dD.as<double>();
  HIPCHK(hipMemcpyAsync(D, Dn, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(D + n, Dss, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(D + 2 * n, Dds, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dn.p, normals, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(flags.p, free_flags, 3 * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(geom.p, geometry, sizeof(double) * 3 * n * CT_GEOM, hipMemcpyHostToDevice, st));
  HIPCHK(stress_intensity(D, D + n, D + 2 * n, n, G, nu, dn.as<double>(), flags.as<uint8_t>(), geom.as<double>(),
                          dK.as<double>(), st));
  HIPCHK(hipMemcpyAsync(K, dK.p, sizeof(double) * 3 * n * CT_K, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

int cdj_stress_intensity_dev(cdj_ctx* ctx, const double* midpoints, const double* P1, const double* P2,
                             const double* P3, const double* normals, int64_t n, const double* Dn, const double* Dss,
                             const double* Dds, double G, double nu, uint8_t* free_flags, double* geometry, double* K,
                             void* stream) {
  if (!ctx) return CDJ_ERR_ARG;
  if (int rc = crack_tip_check(ctx, n)) return rc;
  if (!midpoints || !P1 || !P2 || !P3 || !normals || !Dn || !Dss || !Dds || !free_flags || !geometry || !K)
    return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  if (n == 0) return CDJ_OK;
  DBuf slots, matched;
  HIPCHK(slots.alloc(sizeof(int) * crack_tip_slots(n)));
  HIPCHK(matched.alloc(sizeof(int) * 3 * n));
  HIPCHK(hipEventRecord(ctx->ev0, st));
  HIPCHK(crack_tip_elements(midpoints, P1, P2, P3, normals, n, slots.as<int>(), matched.as<int>(), free_flags,
                            geometry, st));
  HIPCHK(stress_intensity(Dn, Dss, Dds, n, G, nu, normals, free_flags, geometry, K, st));
  HIPCHK(hipEventRecord(ctx->ev1, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

}  // extern "C"

// ---- frictional slip (SlipCalculator3D.jl:36-233) ----------------------------------------------------------
extern "C" {

int cdj_slip_friction(cdj_ctx* ctx, const double* P1, const double* P2, const double* P3, int64_t n, double nu,
                      double G, double lambda, const double* midpoints, const double* normals, int32_t hs_flag,
                      const double* b, const double* mu, const double* Sf, double* Dn, double* Dss, double* Dds,
                      cdj_report* report) {
  if (!ctx) return CDJ_ERR_ARG;
  if (n < 0) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  if (!P1 || !P2 || !P3 || !midpoints || !normals || !b || !mu || !Sf || !Dn || !Dss || !Dds)
    return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  if (report) *report = cdj_report{};
  if (n == 0) return CDJ_OK;
  if (5 * n > 65535) return fail(ctx, CDJ_ERR_DIMS, "frictional solve too large (5 n > 65535)");
  // SlipCalculator3D.jl:52-59: no element can overcome its frictional strength -> no slip
  bool locked = true;
  for (int64_t i = 0; i < n && locked; ++i) {
    const double fr = b[i] * mu[i];
    locked = (-std::fabs(b[n + i]) > fr) && (-std::fabs(b[2 * n + i]) > fr);
  }
  if (locked) {
    for (int64_t i = 0; i < n; ++i) Dn[i] = Dss[i] = Dds[i] = 0.0;
    return CDJ_OK;
  }
  // scale: mean half perimeter / G (SlipCalculator3D.jl:66-68, AreaOfTriangle3D.jl)
  double hsum = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double L[3];
    const double* Q[3] = {P1, P2, P3};
    for (int s = 0; s < 3; ++s) {
      const double* A = Q[s];
      const double* B = Q[(s + 1) % 3];
      const double dx = A[i] - B[i], dy = A[n + i] - B[n + i], dz = A[2 * n + i] - B[2 * n + i];
      L[s] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    hsum += (L[0] + L[1] + L[2]) / 2.0;
  }
  const double scale = (hsum / (double)n) / G;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t N3 = 3 * n, N5 = 5 * n;
  const size_t v3 = sizeof(double) * 3 * n;
  DBuf dn, dm, d1, d2, d3, AI, db, dD, dmu, dSf, a, bb, J, x, vec, red, ipiv, info, ws, D3, part, iwork;
  HIPCHK(dn.alloc(v3));
  HIPCHK(dm.alloc(v3));
  HIPCHK(d1.alloc(v3));
  HIPCHK(d2.alloc(v3));
  HIPCHK(d3.alloc(v3));
  HIPCHK(hipMemcpyAsync(dn.p, normals, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dm.p, midpoints, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d1.p, P1, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d2.p, P2, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d3.p, P3, v3, hipMemcpyHostToDevice, st));
  HIPCHK(AI.alloc(sizeof(double) * N3 * 2 * N3));  // [A | I], then [LU | A^-1]
  cdj_opts o{};
  o.precision = CDJ_PREC_FP64;
  o.assembly = CDJ_ASM_FP64;
  int rc = cdj_assemble_dev(ctx, dn.as<double>(), dm.as<double>(), d1.as<double>(), d2.as<double>(), d3.as<double>(),
                            n, nu, G, lambda, hs_flag, AI.as<double>(), N3, &o, st);
  if (rc) return rc;
  scale_matrix(AI.as<double>(), N3 * N3, scale, st);
  HIPCHK(ipiv.alloc(sizeof(int32_t) * N5));
  HIPCHK(info.alloc(sizeof(int)));
  HIPCHK(ws.alloc(lu_workspace_bytes(N5)));
  HIPCHK(invert(AI.as<double>(), N3, ipiv.as<int32_t>(), info.as<int>(), ws.p, st));
  int h_info = 0;
  HIPCHK(hipMemcpyAsync(&h_info, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (h_info < 0) return fail(ctx, CDJ_ERR_DEVICE, "LU panel barrier wait timed out");
  if (h_info > 0) return fail(ctx, CDJ_ERR_SINGULAR, "matrix is exactly singular (zero pivot " + std::to_string(h_info) + ")");
  const double* C = AI.as<double>() + N3 * N3;
  HIPCHK(db.alloc(v3));
  HIPCHK(dD.alloc(v3));
  HIPCHK(part.alloc(std::max(gemv_part_bytes(N5, N5), gemv_part_bytes(N3, N3))));
  HIPCHK(hipMemcpyAsync(db.p, b, v3, hipMemcpyHostToDevice, st));
  gemv_splitk<double>(C, N3, N3, N3, db.as<double>(), dD.as<double>(), 1.0, nullptr, 0.0, part.as<double>(), st);
  HIPCHK(dmu.alloc(sizeof(double) * n));
  HIPCHK(dSf.alloc(sizeof(double) * n));
  HIPCHK(hipMemcpyAsync(dmu.p, mu, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dSf.p, Sf, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(a.alloc(sizeof(double) * N5 * N5));
  HIPCHK(bb.alloc(sizeof(double) * N5));
  build_friction_lcp(C, n, dD.as<double>(), dmu.as<double>(), dSf.as<double>(), a.as<double>(), bb.as<double>(), st);
  HIPCHK(hipGetLastError());
  AI.reset();  // the inverse is folded into a: free its 18 n^2 doubles before the Newton scratch
  // frictionless elements (mu = Sf = 0) leave degenerate unknowns out of the Newton solve (lcp_keep)
  const std::vector<int32_t> keep = lcp_keep(mu, Sf, n);
  const int64_t Nr = (int64_t)keep.size();
  DBuf dkeep, ar, br, xr;
  if (Nr < N5) {
    HIPCHK(dkeep.alloc(sizeof(int32_t) * Nr));
    HIPCHK(ar.alloc(sizeof(double) * Nr * Nr));
    HIPCHK(br.alloc(sizeof(double) * Nr));
    HIPCHK(xr.alloc(sizeof(double) * Nr));
    HIPCHK(hipMemcpyAsync(dkeep.p, keep.data(), sizeof(int32_t) * Nr, hipMemcpyHostToDevice, st));
    lcp_gather(a.as<double>(), bb.as<double>(), N5, dkeep.as<int32_t>(), Nr, ar.as<double>(), br.as<double>(), st);
  }
  HIPCHK(J.alloc(sizeof(double) * Nr * Nr));
  HIPCHK(x.alloc(sizeof(double) * N5));
  HIPCHK(vec.alloc(sizeof(double) * 8 * N5));
  HIPCHK(red.alloc(sizeof(double) * 8));
  FnParams fp;
  FnResult fr;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  HIPCHK(iwork.alloc(sizeof(int) * fn_iwork_ints(Nr)));
  const bool red_lcp = Nr < N5;
  const double* la = red_lcp ? ar.as<double>() : a.as<double>();
  const double* lb = red_lcp ? br.as<double>() : bb.as<double>();
  double* lx = red_lcp ? xr.as<double>() : x.as<double>();
  hipError_t fe = fischer_newton(la, lb, Nr, lx, fp, fr, J.as<double>(), ipiv.as<int32_t>(), info.as<int>(), ws.p,
                                 vec.as<double>(), red.as<double>(), part.as<double>(), iwork.as<int>(), st);
  if (fe != hipSuccess) return hip_fail(ctx, fe, "Fischer-Newton");
  if (red_lcp) lcp_scatter(xr.as<double>(), dkeep.as<int32_t>(), Nr, x.as<double>(), N5, st);
  HIPCHK(D3.alloc(v3));
  friction_slip(a.as<double>(), bb.as<double>(), x.as<double>(), n, scale, vec.as<double>(), D3.as<double>(),
                part.as<double>(), st);
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipMemcpyAsync(Dn, D3.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(Dss, D3.as<double>() + n, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(Dds, D3.as<double>() + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (report) {
    report->path = CDJ_PREC_FP64;
    report->iterations = fr.iterations;
    report->residual = fr.err;
    report->solve_ms = ms;
  }
  return CDJ_OK;
}

}  // extern "C"

// ---- constant fluid volume (SlipCalculator3D.jl:242-475, ComputePressurisedCrackDn.jl) ----------------------

namespace {

// Area and half perimeter of each triangle (AreaOfTriangle3D.jl: Heron's formula on the edge lengths)
void area_half_perimeter(const double* P1, const double* P2, const double* P3, int64_t n, std::vector<double>& area,
                         std::vector<double>& hp) {
  area.assign(n, 0.0);
  hp.assign(n, 0.0);
  const double* Q[3] = {P1, P2, P3};
  for (int64_t i = 0; i < n; ++i) {
    double L[3];
    for (int s = 0; s < 3; ++s) {
      const double* A = Q[s];
      const double* B = Q[(s + 1) % 3];
      const double dx = A[i] - B[i], dy = A[n + i] - B[n + i], dz = A[2 * n + i] - B[2 * n + i];
      L[s] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    const double h = (L[0] + L[1] + L[2]) / 2.0;
    hp[i] = h;
    area[i] = std::sqrt(h * (h - L[0]) * (h - L[1]) * (h - L[2]));
  }
}

}  // namespace

extern "C" {

int cdj_slip_fluid_volume(cdj_ctx* ctx, const double* P1, const double* P2, const double* P3, int64_t n, double nu,
                          double G, double lambda, const double* midpoints, const double* normals, int32_t hs_flag,
                          const double* b, const int32_t* fracture_flag, double volume, double* Dn, double* Dss,
                          double* Dds, double* pressure, cdj_report* report) {
  if (!ctx) return CDJ_ERR_ARG;
  if (n < 0) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  if (!P1 || !P2 || !P3 || !midpoints || !normals || !b || !fracture_flag || !Dn || !Dss || !Dds)
    return fail(ctx, CDJ_ERR_ARG, "NULL argument");
  if (report) *report = cdj_report{};
  if (pressure) *pressure = 0.0;
  if (n == 0) return CDJ_OK;
  if (5 * n > 65535) return fail(ctx, CDJ_ERR_DIMS, "frictional solve too large (5 n > 65535)");
  // fracture 1 is the crack (SlipCalculator3D.jl:286: more than one fracture ends in the reference's error),
  // 0 marks free-surface elements (ComputePressurisedCrackDn.jl:43-46)
  bool any_free = false;
  for (int64_t i = 0; i < n; ++i) {
    if (fracture_flag[i] > 1) return fail(ctx, CDJ_ERR_ARG, "Check scl here - what do we use for one crack below");
    any_free = any_free || fracture_flag[i] == 0;
  }
  std::vector<double> area, hp;
  area_half_perimeter(P1, P2, P3, n, area, hp);
  double hsum = 0.0, asum = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    hsum += hp[i];
    asum += area[i];
  }
  const double scale = (hsum / (double)n) / G;  // SlipCalculator3D.jl:257-258
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  const int64_t N3 = 3 * n, N5 = 5 * n;
  const size_t v3 = sizeof(double) * N3;
  DBuf dn, dm, d1, d2, d3, AI, dB, dD, dmu, dSf, a, bb0, bb, shift, J, x, vec, red, ipiv, info, ws, D3, dz, part,
      iwork;
  HIPCHK(dn.alloc(v3));
  HIPCHK(dm.alloc(v3));
  HIPCHK(d1.alloc(v3));
  HIPCHK(d2.alloc(v3));
  HIPCHK(d3.alloc(v3));
  HIPCHK(hipMemcpyAsync(dn.p, normals, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dm.p, midpoints, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d1.p, P1, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d2.p, P2, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d3.p, P3, v3, hipMemcpyHostToDevice, st));
  HIPCHK(AI.alloc(sizeof(double) * N3 * 2 * N3));  // [A | I], then [LU | A^-1]
  cdj_opts o{};
  o.precision = CDJ_PREC_FP64;
  o.assembly = CDJ_ASM_FP64;
  int rc = cdj_assemble_dev(ctx, dn.as<double>(), dm.as<double>(), d1.as<double>(), d2.as<double>(), d3.as<double>(),
                            n, nu, G, lambda, hs_flag, AI.as<double>(), N3, &o, st);
  if (rc) return rc;
  scale_matrix(AI.as<double>(), N3 * N3, scale, st);
  // Norm: the largest normal traction of a uniform opening of the target volume (SlipCalculator3D.jl:348-356)
  std::vector<double> h3(N3, 0.0);
  const double height = (volume / asum) / scale;
  for (int64_t i = 0; i < n; ++i) h3[i] = height;
  HIPCHK(dB.alloc(v3));
  HIPCHK(dD.alloc(v3));
  HIPCHK(part.alloc(std::max(gemv_part_bytes(N5, N5), gemv_part_bytes(N3, N3))));
  HIPCHK(hipMemcpyAsync(dB.p, h3.data(), v3, hipMemcpyHostToDevice, st));
  gemv_splitk<double>(AI.as<double>(), N3, N3, N3, dB.as<double>(), dD.as<double>(), 1.0, nullptr, 0.0,
                      part.as<double>(), st);
  HIPCHK(hipMemcpyAsync(h3.data(), dD.p, v3, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double norm = h3[0];
  for (int64_t i = 1; i < n; ++i) norm = std::max(norm, h3[i]);
  HIPCHK(ipiv.alloc(sizeof(int32_t) * N5));
  HIPCHK(info.alloc(sizeof(int)));
  HIPCHK(ws.alloc(lu_workspace_bytes(N5)));
  HIPCHK(invert(AI.as<double>(), N3, ipiv.as<int32_t>(), info.as<int>(), ws.p, st));
  int h_info = 0;
  HIPCHK(hipMemcpyAsync(&h_info, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (h_info < 0) return fail(ctx, CDJ_ERR_DEVICE, "LU panel barrier wait timed out");
  if (h_info > 0) return fail(ctx, CDJ_ERR_SINGULAR, "matrix is exactly singular (zero pivot " + std::to_string(h_info) + ")");
  const double* C = AI.as<double>() + N3 * N3;
  // the frictional problem with the reference's fixed parameters mu = 0.1, Sf = 0 (SlipCalculator3D.jl:275-276)
  const std::vector<double> mu(n, 0.1), Sf(n, 0.0), zeros(N3, 0.0);
  HIPCHK(dmu.alloc(sizeof(double) * n));
  HIPCHK(dSf.alloc(sizeof(double) * n));
  HIPCHK(dz.alloc(v3));
  HIPCHK(hipMemcpyAsync(dmu.p, mu.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dSf.p, Sf.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dz.p, zeros.data(), v3, hipMemcpyHostToDevice, st));
  HIPCHK(a.alloc(sizeof(double) * N5 * N5));
  HIPCHK(bb0.alloc(sizeof(double) * N5));
  HIPCHK(bb.alloc(sizeof(double) * N5));
  build_friction_lcp(C, n, dz.as<double>(), dmu.as<double>(), dSf.as<double>(), a.as<double>(), bb0.as<double>(), st);
  HIPCHK(hipGetLastError());
  // free-surface elements: the shift that keeps their opening entries non-negative (ComputePressurisedCrackDn.jl:
  // 185-193); the value added to every free element's entry, and subtracted from its opening afterwards
  double arb = 0.0;
  if (any_free) {
    std::vector<double> h0(n);
    HIPCHK(hipMemcpyAsync(h0.data(), bb0.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double mn = INFINITY;
    for (int64_t i = 0; i < n; ++i)
      if (fracture_flag[i] == 0) mn = std::min(mn, h0[i]);
    arb = std::fabs(mn * 2);
  }
  std::vector<double> hshift(N5, 0.0);
  for (int64_t i = 0; i < n; ++i)
    if (fracture_flag[i] == 0) hshift[i] = arb;
  HIPCHK(shift.alloc(sizeof(double) * N5));
  HIPCHK(hipMemcpyAsync(shift.p, hshift.data(), sizeof(double) * N5, hipMemcpyHostToDevice, st));
  HIPCHK(J.alloc(sizeof(double) * N5 * N5));
  HIPCHK(x.alloc(sizeof(double) * N5));
  HIPCHK(vec.alloc(sizeof(double) * 8 * N5));
  HIPCHK(red.alloc(sizeof(double) * 8));
  HIPCHK(D3.alloc(v3));
  HIPCHK(iwork.alloc(sizeof(int) * fn_iwork_ints(N5)));
  FnParams fp;
  int evaluations = 0, newton = 0;
  hipError_t herr = hipSuccess;
  std::vector<double> hB(N3), hD(N3), s3(N3);
  // ComputePressurisedCrackDn(x::Tractions, ...): slip at pressure x Norm on the fracture elements, into s3
  auto slip_at = [&](double xp) -> bool {
    ++evaluations;
    for (int64_t i = 0; i < N3; ++i) hB[i] = b[i];
    const double add = xp * norm;
    for (int64_t i = 0; i < n; ++i)
      if (fracture_flag[i] == 1) hB[i] += add;
    if ((herr = hipMemcpyAsync(dB.p, hB.data(), v3, hipMemcpyHostToDevice, st)) != hipSuccess) return false;
    gemv_splitk<double>(C, N3, N3, N3, dB.as<double>(), dD.as<double>(), 1.0, nullptr, 0.0, part.as<double>(), st);
    if ((herr = hipMemcpyAsync(hD.data(), dD.p, sizeof(double) * n, hipMemcpyDeviceToHost, st)) != hipSuccess)
      return false;
    if ((herr = hipStreamSynchronize(st)) != hipSuccess) return false;
    bool opening = false;
    for (int64_t i = 0; i < n && !opening; ++i) opening = (!any_free || fracture_flag[i] != 0) && hD[i] > 0;
    if (!opening) {  // already closed: no slip
      std::fill(s3.begin(), s3.end(), 0.0);
      return true;
    }
    // (bb0 + D) + shift where the reference forms (bb0 + shift) + D: with Sf = 0 the shift is zero, so the two agree
    build_friction_lcp_vector(C, n, dD.as<double>(), dSf.as<double>(), bb.as<double>(), st);
    if (any_free) add_vector(bb.as<double>(), shift.as<double>(), N5, st);
    FnResult fr;
    if ((herr = fischer_newton(a.as<double>(), bb.as<double>(), N5, x.as<double>(), fp, fr, J.as<double>(),
                               ipiv.as<int32_t>(), info.as<int>(), ws.p, vec.as<double>(), red.as<double>(),
                               part.as<double>(), iwork.as<int>(), st)) != hipSuccess)
      return false;
    newton += fr.iterations;
    friction_slip(a.as<double>(), bb.as<double>(), x.as<double>(), n, scale, vec.as<double>(), D3.as<double>(),
                  part.as<double>(), st);
    if ((herr = hipMemcpyAsync(s3.data(), D3.p, v3, hipMemcpyDeviceToHost, st)) != hipSuccess) return false;
    if ((herr = hipStreamSynchronize(st)) != hipSuccess) return false;
    if (any_free)
      for (int64_t i = 0; i < n; ++i) s3[i] -= arb;
    return true;
  };
  auto vol_of = [&]() {
    double v = 0.0;
    for (int64_t i = 0; i < n; ++i)
      if (fracture_flag[i] == 1) v += s3[i] * area[i];
    return v;
  };
  auto volume_at = [&](double xp, double& v) -> bool {
    if (!slip_at(xp)) return false;
    v = vol_of();
    return true;
  };
  auto misfit = [&](double xp, double& f) -> bool {
    double v;
    if (!volume_at(xp, v)) return false;
    f = std::fabs(volume - v) * std::fabs(volume - v);
    return true;
  };
  auto dev_fail = [&]() { return hip_fail(ctx, herr == hipSuccess ? hipErrorUnknown : herr, "fluid-volume solve"); };
  // scale Norm until the volume at x = 10 lies in [V, 2.5 V] (SlipCalculator3D.jl:365-392)
  constexpr int kNormLoops = 200;
  double huge = 0.0;
  if (!volume_at(10.0, huge)) return dev_fail();
  for (int it = 0; huge < volume; ++it) {
    if (it >= kNormLoops) return fail(ctx, CDJ_ERR_ARG, "the target volume is not reached at any pressure scale");
    norm *= 1.05;
    if (!volume_at(10.0, huge)) return dev_fail();
  }
  double reduce = 0.3;
  for (int it = 0; huge > volume * 2.5; ++it) {
    if (it >= kNormLoops) return fail(ctx, CDJ_ERR_ARG, "the pressure scale does not settle");
    norm *= (1 - reduce);
    if (!volume_at(10.0, huge)) return dev_fail();
    if (huge < volume) {
      norm /= (1 - reduce);
      reduce /= 2;
      if (!volume_at(10.0, huge)) return dev_fail();
    }
  }
  // Brent on [0, 10] with abs_tol 0.001; golden section on [0, 50] when its volume is negative
  // (SlipCalculator3D.jl:404-431)
  Minimum mn = brent_minimize(misfit, 0.0, 10.0, 0.001);
  if (!mn.ok) return dev_fail();
  double vol_out = 0.0;
  if (!volume_at(mn.x, vol_out)) return dev_fail();
  if ((volume - vol_out) > volume) {
    mn = golden_section_minimize(misfit, 0.0, 50.0);
    if (!mn.ok) return dev_fail();
  }
  if (!slip_at(mn.x)) return dev_fail();
  const double vol = vol_of();
  const bool bad = std::fabs(vol) > std::fabs(volume * 2);  // SlipCalculator3D.jl:460-465: NaN the result
  for (int64_t i = 0; i < n; ++i) {
    Dn[i] = bad ? NAN : s3[i];
    Dss[i] = bad ? NAN : s3[n + i];
    Dds[i] = bad ? NAN : s3[2 * n + i];
  }
  if (pressure) *pressure = mn.x * norm;
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (report) {
    report->path = CDJ_PREC_FP64;
    report->iterations = evaluations;
    report->residual = volume != 0.0 ? std::fabs(volume - vol) / std::fabs(volume) : std::fabs(vol);
    report->solve_ms = ms;
    report->near_entries = newton;
  }
  return CDJ_OK;
}

}  // extern "C"

// ---- MLP evaluation (the learned preconditioner's network on one mesh) -------------------------------------
extern "C" {

int cdj_mlp_forward(cdj_ctx* ctx, const cdj_mlp* mlp, const double* P1, const double* P2, const double* P3, int64_t n,
                    int32_t bf16, float* Y) {
  if (!ctx || !mlp || !P1 || !P2 || !P3 || !Y) return CDJ_ERR_ARG;
  if (n <= 0) return fail(ctx, CDJ_ERR_DIMS, "Element size inputs must be the same dimensions");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const Mlp& m = mlp->m;
  const size_t v3 = sizeof(double) * 3 * n;
  DBuf d1, d2, d3, acts, wbf;
  HIPCHK(d1.alloc(v3));
  HIPCHK(d2.alloc(v3));
  HIPCHK(d3.alloc(v3));
  HIPCHK(hipMemcpyAsync(d1.p, P1, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d2.p, P2, v3, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d3.p, P3, v3, hipMemcpyHostToDevice, st));
  MeshGuard g;
  HIPCHK(mesh_prepare(d1.as<double>(), d2.as<double>(), d3.as<double>(), n, m.k, g.f, st));
  HIPCHK(acts.alloc(sizeof(float) * n * (2 * m.H + m.O)));
  if (bf16) {
    HIPCHK(wbf.alloc(sizeof(unsigned short) * mlp_bf16_weight_count(m)));
    HIPCHK(mlp_forward_bf16(m, g.f.feat, n, acts.as<float>(), wbf.as<unsigned short>(), st));
  } else {
    HIPCHK(mlp_forward(m, g.f.feat, n, acts.as<float>(), st));
  }
  HIPCHK(hipMemcpyAsync(Y, acts.as<float>() + 2 * n * m.H, sizeof(float) * n * m.O, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return CDJ_OK;
}

}  // extern "C"
