// C ABI of include/huv.h: host engine entry points (FP64) and the device
// context / kernel-set / batched SOCS image entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/huv.h"
#include "capi_internal.hpp"
#include "hip/mask_kernels.hpp"
#include "hip/socs_kernels.hpp"
#include "host/engine.hpp"

namespace {
thread_local std::string g_err;
}

// Device scratch of the no-allocation entry point (huv_socs_images_noalloc):
// one copy per GPU, held until the process ends.  Host code never uses this
// name as an address: each device's copy is resolved once with
// hipGetSymbolAddress (scratch_address below) and that pointer is what the
// kernels receive.
constexpr size_t kScratchBytes = (size_t)HUV_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, char*> g_scratch_ptr;  // device -> its copy of `scratch`

// The device's copy of `scratch`; `device` must be current.
hipError_t scratch_address(int device, char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return hipSuccess;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return e != hipSuccess ? e : hipErrorInvalidSymbol;
  g_scratch_ptr[device] = (char*)p;
  *out = (char*)p;
  return hipSuccess;
}
}  // namespace

namespace huvi {

huv_status fail(huv_status code, const std::string& msg) {
  g_err = msg;
  return code;
}

huv_status hip_fail(hipError_t e, const char* what) {
  return fail(HUV_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

void* Workspace::get(int slot, size_t bytes, hipError_t& err) {
  err = hipSuccess;
  if (bytes == 0) bytes = 256;
  if (size[slot] >= bytes) return ptr[slot];
  if (ptr[slot]) {
    // Work already queued may still use the old buffer.
    err = hipDeviceSynchronize();
    if (err != hipSuccess) return nullptr;
    err = hipFree(ptr[slot]);
    ptr[slot] = nullptr;
    size[slot] = 0;
    if (err != hipSuccess) return nullptr;
  }
  err = hipMalloc(&ptr[slot], bytes);
  ++generation;
  if (err != hipSuccess) {
    ptr[slot] = nullptr;
    return nullptr;
  }
  size[slot] = bytes;
  return ptr[slot];
}

void Workspace::release() {
  for (int i = 0; i < kSlots; ++i) {
    if (ptr[i]) hipFree(ptr[i]);
    ptr[i] = nullptr;
    size[i] = 0;
  }
}

void* Ctx::buf(int slot, size_t bytes, hipError_t& err) {
  if (!fixed_mode) return ws.get(slot, bytes, err);
  err = hipSuccess;
  const size_t need = (bytes + 255) / 256 * 256;
  if (fixed_used + need > fixed_cap) {
    err = hipErrorOutOfMemory;
    return nullptr;
  }
  void* p = fixed + fixed_used;
  fixed_used += need;
  return p;
}

SocsSizes socs_sizes(int nclips, int n, int B, int m, bool paired, int nplanes, int S,
                     bool raster) {
  auto r = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t W = 2 * (size_t)huvk::coarse_band(m) + 1;
  const size_t nimg = (size_t)nclips * nplanes;
  const size_t G = 2 * (size_t)B;
  SocsSizes z;
  if (raster) z.tmp_rows = r((size_t)nclips * (2 * B + 1) * (n / 2) * sizeof(float2));
  z.band = r((size_t)nclips * W * W * sizeof(float2));
  if (paired) z.band_paired = r((size_t)nclips * (W + 1) * W * sizeof(float2));
  z.coarse = r((size_t)S * nimg * m * m * sizeof(float));
  z.tmp_coarse = r(nimg * (2 * G + 1) * (m / 2) * sizeof(float2));
  z.iband = r(nimg * (G + 1) * (2 * G + 1) * sizeof(float2));
  return z;
}

const float2* Ctx::twiddles(int N, hipError_t& err) {
  err = hipSuccess;
  auto it = tw.find(N);
  if (it != tw.end()) return it->second;
  if (fixed_mode) {  // no-allocation calls only read tables built earlier
    err = hipErrorNotReady;
    return nullptr;
  }
  std::vector<float2> h((size_t)N);
  for (int k = 0; k < N; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)N;
    h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  float2* d = nullptr;
  err = hipMalloc(&d, sizeof(float2) * N);
  if (err != hipSuccess) return nullptr;
  err = hipMemcpy(d, h.data(), sizeof(float2) * N, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    hipFree(d);
    return nullptr;
  }
  tw[N] = d;
  return d;
}

static long sinc_key(int n, int B, bool pixel_transfer) {
  return ((long)n << 20) | ((long)B << 1) | (pixel_transfer ? 1 : 0);
}

const float* Ctx::sinc_table(int n, int B, bool pixel_transfer, hipError_t& err) {
  err = hipSuccess;
  const long key = sinc_key(n, B, pixel_transfer);
  auto it = sinc.find(key);
  if (it != sinc.end()) return it->second;
  if (fixed_mode) {
    err = hipErrorNotReady;
    return nullptr;
  }
  const int W = 2 * B + 1;
  std::vector<float> h((size_t)W);
  for (int k = -B; k <= B; ++k) {
    double v = 1.0;
    if (pixel_transfer) {
      const double u = M_PI * (double)k / (double)n;
      v = std::fabs(u) < 1e-6 ? 1.0 - u * u / 6.0 : std::sin(u) / u;
    }
    h[k + B] = (float)v;
  }
  float* d = nullptr;
  err = hipMalloc(&d, sizeof(float) * W);
  if (err != hipSuccess) return nullptr;
  err = hipMemcpy(d, h.data(), sizeof(float) * W, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    hipFree(d);
    return nullptr;
  }
  sinc[key] = d;
  return d;
}

void Ctx::stage_begin(int stage, hipEvent_t* a) {
  (void)stage;
  *a = nullptr;
  if (!profiling) return;
  if (hipEventCreate(a) != hipSuccess) { *a = nullptr; return; }
  hipEventRecord(*a, stream());
}

void Ctx::stage_end(int stage, hipEvent_t a) {
  if (!profiling || !a) return;
  hipEvent_t b;
  if (hipEventCreate(&b) != hipSuccess) return;
  hipEventRecord(b, stream());
  pending.push_back({stage, a, b});
}

void Ctx::collect_stage_times() {
  for (auto& p : pending) {
    hipEventSynchronize(p.b);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      stage_ms[p.stage] += ms;
      stage_calls[p.stage] += 1;
    }
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  pending.clear();
}

Ctx::~Ctx() {
  hipSetDevice(device);
  if (own_stream) hipStreamSynchronize(own_stream);
  collect_stage_times();
  ws.release();
  for (auto& kv : tw) hipFree(kv.second);
  for (auto& kv : sinc) hipFree(kv.second);
  for (auto& kv : plane_tables) hipFree(kv.second);
  for (auto& kv : plane_packs) {
    hipFree(kv.second.kern);
    hipFree(kv.second.lam);
  }
  for (auto& kv : graphs) drop_graph(kv.second);
  if (pin_h) hipHostFree(pin_h);
  if (pin_d) hipFree(pin_d);
  if (dl_h) hipHostFree(dl_h);
  if (own_stream) hipStreamDestroy(own_stream);
}

bool Ctx::staging(size_t bytes, hipError_t& e) {
  e = hipSuccess;
  if (bytes <= pin_cap) return true;
  e = hipStreamSynchronize(stream());
  if (e != hipSuccess) return false;
  if (pin_h) hipHostFree(pin_h);
  if (pin_d) hipFree(pin_d);
  pin_h = pin_d = nullptr;
  pin_cap = 0;
  e = hipHostMalloc((void**)&pin_h, bytes, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void**)&pin_d, bytes);
  if (e != hipSuccess) {
    if (pin_h) hipHostFree(pin_h);
    pin_h = nullptr;
    pin_d = nullptr;
    return false;
  }
  pin_cap = bytes;
  return true;
}

bool Ctx::download_area(size_t bytes, hipError_t& e) {
  e = hipSuccess;
  if (bytes <= dl_cap) return true;
  e = hipStreamSynchronize(stream());
  if (e != hipSuccess) return false;
  if (dl_h) hipHostFree(dl_h);
  dl_h = dl_d = nullptr;
  dl_cap = 0;
  e = hipHostMalloc((void**)&dl_h, bytes, hipHostMallocMapped);
  if (e != hipSuccess) {
    dl_h = nullptr;
    return false;
  }
  e = hipHostGetDevicePointer((void**)&dl_d, dl_h, 0);
  if (e != hipSuccess) {
    hipHostFree(dl_h);
    dl_h = dl_d = nullptr;
    return false;
  }
  dl_cap = bytes;
  return true;
}

void Ctx::drop_graph(GraphEntry& g) {
  if (g.exec) hipGraphExecDestroy(g.exec);
  if (g.graph) hipGraphDestroy(g.graph);
  g.exec = nullptr;
  g.graph = nullptr;
}

void Ctx::forget_kset(uint64_t id) {
  bool busy = false;
  for (auto it = graphs.begin(); it != graphs.end();) {
    if (std::find(it->second.ksets.begin(), it->second.ksets.end(), id) != it->second.ksets.end()) {
      if (!busy) hipStreamSynchronize(stream());
      if (!busy && own_stream) hipStreamSynchronize(own_stream);
      busy = true;
      drop_graph(it->second);
      it = graphs.erase(it);
    } else {
      ++it;
    }
  }
  for (auto it = plane_tables.begin(); it != plane_tables.end();) {
    if (std::find(it->first.begin(), it->first.end(), id) != it->first.end()) {
      if (!busy) hipStreamSynchronize(stream());
      busy = true;
      hipFree(it->second);
      it = plane_tables.erase(it);
    } else {
      ++it;
    }
  }
  for (auto it = plane_packs.begin(); it != plane_packs.end();) {
    if (std::find(it->first.begin(), it->first.end(), id) != it->first.end()) {
      if (!busy) hipStreamSynchronize(stream());
      busy = true;
      hipFree(it->second.kern);
      hipFree(it->second.lam);
      it = plane_packs.erase(it);
    } else {
      ++it;
    }
  }
}

huv_status band_stage(Ctx* c, int nclips, int n, int B, int BB, const float* masks, float2* Mq,
                      bool pixel_transfer, int n1, float2* Mp) {
  hipError_t e;
  const float2* tw_n = c->twiddles(n, e);
  if (!tw_n) return hip_fail(e, "twiddle table");
  const float* st = c->sinc_table(n, B, pixel_transfer, e);
  if (!st) return hip_fail(e, "sinc table");
  const size_t tmp_bytes = (size_t)nclips * (2 * B + 1) * (n / 2) * sizeof(float2);
  float2* tmp = (float2*)c->buf(Workspace::TMP_ROWS, tmp_bytes, e);
  if (!tmp) return hip_fail(e, "workspace");
  hipEvent_t ev;
  c->stage_begin(ST_ROWS_BAND, &ev);
  e = huvk::launch_rows_band(n, n1, B, nclips, masks, tw_n, tmp, c->stream());
  if (e != hipSuccess) return hip_fail(e, "rows_band launch");
  c->stage_end(ST_ROWS_BAND, ev);
  c->stage_begin(ST_COLS_BAND, &ev);
  e = huvk::launch_cols_band(n, n1, B, BB, nclips, tmp, tw_n, st, Mq, Mp, c->stream());
  if (e != hipSuccess) return hip_fail(e, "cols_band launch");
  c->stage_end(ST_COLS_BAND, ev);
  return HUV_OK;
}

huv_status band_stage_complex(Ctx* c, int nclips, int n, int B, int BB, const float2* masks,
                              float2* Mq, bool pixel_transfer, int n1, float2* Mp) {
  hipError_t e;
  const float2* tw_n = c->twiddles(n, e);
  if (!tw_n) return hip_fail(e, "twiddle table");
  const float* st = c->sinc_table(n, B, pixel_transfer, e);
  if (!st) return hip_fail(e, "sinc table");
  const size_t tmp_bytes = (size_t)nclips * (2 * B + 1) * n * sizeof(float2);
  float2* tmp = (float2*)c->buf(Workspace::TMP_ROWS, tmp_bytes, e);
  if (!tmp) return hip_fail(e, "workspace");
  hipEvent_t ev;
  c->stage_begin(ST_ROWS_BAND, &ev);
  e = huvk::launch_rows_band_c(n, n1, B, nclips, masks, tw_n, tmp, c->stream());
  if (e != hipSuccess) return hip_fail(e, "rows_band_c launch");
  c->stage_end(ST_ROWS_BAND, ev);
  c->stage_begin(ST_COLS_BAND, &ev);
  e = huvk::launch_cols_band_c(n, n1, B, BB, nclips, tmp, tw_n, st, Mq, Mp, c->stream());
  if (e != hipSuccess) return hip_fail(e, "cols_band_c launch");
  c->stage_end(ST_COLS_BAND, ev);
  return HUV_OK;
}

huv_status spectrum_band_stage(Ctx* c, int nclips, int n, int B, int BB, const double* spectra,
                               float2* Mq, float2* Mp) {
  const int W = 2 * BB + 1;
  const size_t box = (size_t)W * W, pbox = (size_t)(W + 1) * W;
  const size_t bytes = (size_t)nclips * (box + (Mp ? pbox : 0)) * sizeof(float2);
  hipError_t e = hipStreamSynchronize(c->stream());  // the staging area may feed an earlier copy
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  if (!c->staging(bytes, e)) return hip_fail(e, "staging buffer");
  float2* hq = (float2*)c->pin_h;
  float2* hp = hq + (size_t)nclips * box;
  std::memset(c->pin_h, 0, bytes);
  for (int i = 0; i < nclips; ++i) {
    const double* sp = spectra + (size_t)i * n * n * 2;
    for (int kx = -B; kx <= B; ++kx)
      for (int ky = -B; ky <= B; ++ky) {
        const size_t idx = (size_t)((ky + n) % n) * n + (size_t)((kx + n) % n);
        const float2 v = make_float2((float)sp[2 * idx], (float)sp[2 * idx + 1]);
        const int col = kx + BB, row = ky + BB;
        hq[(size_t)i * box + (size_t)col * W + row] = v;
        if (Mp) hp[(size_t)i * pbox + ((size_t)((col + 1) >> 1) * W + row) * 2 + ((col + 1) & 1)] = v;
      }
  }
  hipEvent_t ev;
  c->stage_begin(ST_COLS_BAND, &ev);
  e = hipMemcpyAsync(Mq, hq, (size_t)nclips * box * sizeof(float2), hipMemcpyHostToDevice,
                     c->stream());
  if (e == hipSuccess && Mp)
    e = hipMemcpyAsync(Mp, hp, (size_t)nclips * pbox * sizeof(float2), hipMemcpyHostToDevice,
                       c->stream());
  if (e != hipSuccess) return hip_fail(e, "spectrum upload");
  c->stage_end(ST_COLS_BAND, ev);
  return HUV_OK;
}


// Bounds on one call's rectangle geometry: the edge grid of one clip (as in
// mask.rs:95, MAX_CELLS) and the host staging of all clips.
constexpr int64_t kRectMaxGrid = 1 << 20;
constexpr int64_t kRectMaxCells = 1ll << 26;
// Polygons: the crossing test is quadratic in the vertex count, so one polygon
// holds at most kPolyMaxVertices vertices; one call stages at most
// kPolyMaxEdges clipped edges.
constexpr int64_t kPolyMaxVertices = 8192;
constexpr int64_t kPolyMaxEdges = 1ll << 26;

huv_status validate_rects(int nclips, int n, const RectBatch& rb) {
  if (!rb.offsets || !rb.background) return fail(HUV_INVALID_PARAMETER, "null rectangle argument");
  if (!(rb.pixel_nm > 0.0) || !std::isfinite(rb.pixel_nm) || n < 1)
    return fail(HUV_INVALID_PARAMETER, "need pixel_nm > 0 and n >= 1");
  if (rb.offsets[0] != 0) return fail(HUV_INVALID_PARAMETER, "rectangle offsets must start at 0");
  for (int i = 0; i < nclips; ++i)
    if (rb.offsets[i + 1] < rb.offsets[i])
      return fail(HUV_INVALID_PARAMETER, "rectangle offsets must be non-decreasing");
  if (rb.offsets[nclips] > 0 && !rb.rects) return fail(HUV_INVALID_PARAMETER, "null rectangles");
  for (int64_t r = 0; r < rb.offsets[nclips]; ++r)
    for (int k = 0; k < 6; ++k)
      if (!std::isfinite(rb.rects[6 * r + k]))
        return fail(HUV_INVALID_PARAMETER, "rectangle values must be finite");
  for (int i = 0; i < 2 * nclips; ++i)
    if (!std::isfinite(rb.background[i]))
      return fail(HUV_INVALID_PARAMETER, "background values must be finite");
  if (rb.polys) {
    const huv_polygons& P = *rb.polys;
    if (!P.poly_offsets) return fail(HUV_INVALID_PARAMETER, "null polygon offsets");
    if (P.poly_offsets[0] != 0) return fail(HUV_INVALID_PARAMETER, "polygon offsets must start at 0");
    for (int i = 0; i < nclips; ++i)
      if (P.poly_offsets[i + 1] < P.poly_offsets[i])
        return fail(HUV_INVALID_PARAMETER, "polygon offsets must be non-decreasing");
    const int64_t np = P.poly_offsets[nclips];
    if (np > 0 && (!P.vertex_offsets || !P.values))
      return fail(HUV_INVALID_PARAMETER, "null polygon arrays");
    if (np > 0 && P.vertex_offsets[0] != 0)
      return fail(HUV_INVALID_PARAMETER, "vertex offsets must start at 0");
    for (int64_t j = 0; j < np; ++j) {
      if (P.vertex_offsets[j + 1] < P.vertex_offsets[j])
        return fail(HUV_INVALID_PARAMETER, "vertex offsets must be non-decreasing");
      if (P.vertex_offsets[j + 1] - P.vertex_offsets[j] > kPolyMaxVertices)
        return fail(HUV_INVALID_PARAMETER, "a polygon has more than 8192 vertices");
    }
    const int64_t nv = np > 0 ? P.vertex_offsets[np] : 0;
    if (nv > 0 && !P.vertices) return fail(HUV_INVALID_PARAMETER, "null polygon vertices");
    for (int64_t k = 0; k < 2 * nv; ++k)
      if (!std::isfinite(P.vertices[k]))
        return fail(HUV_INVALID_PARAMETER, "polygon vertices must be finite");
    for (int64_t k = 0; k < 2 * np; ++k)
      if (!std::isfinite(P.values[k]))
        return fail(HUV_INVALID_PARAMETER, "polygon values must be finite");
  }
  return HUV_OK;
}

namespace {
using Pt = std::pair<double, double>;
double orient2(const Pt& a, const Pt& b, const Pt& c) {
  return (b.first - a.first) * (c.second - a.second) - (b.second - a.second) * (c.first - a.first);
}
// No two non-adjacent edges cross properly (mask.rs:1222-1243).
bool polygon_simple(const std::vector<Pt>& v) {
  const size_t n = v.size();
  for (size_t a = 0; a < n; ++a)
    for (size_t b = a + 2; b < n; ++b) {
      if (a == 0 && b == n - 1) continue;
      const Pt &a0 = v[a], &a1 = v[(a + 1) % n], &b0 = v[b], &b1 = v[(b + 1) % n];
      if (orient2(a0, a1, b0) * orient2(a0, a1, b1) < 0.0 &&
          orient2(b0, b1, a0) * orient2(b0, b1, a1) < 0.0)
        return false;
    }
  return true;
}
// Sutherland-Hodgman clip to [-half, half]^2 (mask.rs:1244-1275); slivers
// along the border of a clipped concave polygon cancel in the edge sum.
std::vector<Pt> clip_polygon(std::vector<Pt> poly, double half) {
  for (int q = 0; q < 4 && poly.size() >= 3; ++q) {
    const bool xa = q < 2;
    const double sg = (q % 2 == 0) ? -1.0 : 1.0, bound = sg * half;
    auto co = [&](const Pt& p) { return xa ? p.first : p.second; };
    std::vector<Pt> o;
    const size_t len = poly.size();
    for (size_t i = 0; i < len; ++i) {
      const Pt cur = poly[i], prev = poly[(i + len - 1) % len];
      const bool pin = sg * co(prev) <= half, cin = sg * co(cur) <= half;
      auto cut = [&]() {
        const double t = (bound - co(prev)) / (co(cur) - co(prev));
        return Pt(prev.first + t * (cur.first - prev.first), prev.second + t * (cur.second - prev.second));
      };
      if (pin && cin) o.push_back(cur);
      else if (pin) o.push_back(cut());
      else if (cin) { o.push_back(cut()); o.push_back(cur); }
    }
    poly.swap(o);
  }
  if (poly.size() < 3) poly.clear();
  return poly;
}
}  // namespace

huv_status rect_band_stage(Ctx* c, int nclips, int n, int B, int BB, const RectBatch& rb,
                           float2* Mq, float2* Mp) {
  // Host plan: disjoint cells per clip (painter's rule), staged with the cell
  // offsets and backgrounds in one pinned upload.
  std::vector<double> cells, edges, pdc((size_t)nclips * 2, 0.0);
  std::vector<int64_t> coff((size_t)nclips + 1, 0), eoff((size_t)nclips + 1, 0);
  std::vector<huv::RectFeature> fs;
  std::vector<huv::RectCell> cl;
  const double half = n * rb.pixel_nm / 2.0;
  // overlaps of at most 1e-9 of the field length count as none (mask.rs:838-841,1099-1108)
  const double eps = 1e-9 * (2.0 * half);
  if (rb.polys && B > huvk::kPolyMaxBand)
    return fail(HUV_UNSUPPORTED, "polygon masks support band half-widths up to " +
                                     std::to_string(huvk::kPolyMaxBand));
  for (int i = 0; i < nclips; ++i) {
    fs.clear();
    for (int64_t r = rb.offsets[i]; r < rb.offsets[i + 1]; ++r) {
      const double* q = rb.rects + 6 * r;
      fs.push_back({q[0], q[1], q[2], q[3], huv::cd(q[4], q[5])});
    }
    const huv::cd bg(rb.background[2 * i], rb.background[2 * i + 1]);
    if (!huv::resolve_rect_cells(fs, bg, n, rb.pixel_nm, kRectMaxGrid, cl))
      return fail(HUV_INVALID_PARAMETER,
                  "clip " + std::to_string(i) + ": more than 2^20 edge-grid cells; rasterise it instead");
    for (const auto& e : cl) cells.insert(cells.end(), {e.x0, e.x1, e.y0, e.y1, e.w.real(), e.w.imag()});
    coff[i + 1] = (int64_t)(cells.size() / 6);
    if (coff[i + 1] > kRectMaxCells) return fail(HUV_INVALID_PARAMETER, "too many rectangle cells in one call");
    if (rb.polys) {
      // bounding boxes of the clip's features, for the overlap rule
      std::vector<std::array<double, 4>> boxes;
      for (const auto& f : fs) {
        if (!(f.w > 0 && f.h > 0)) continue;
        const double x0 = std::max(f.x - f.w / 2, -half), x1 = std::min(f.x + f.w / 2, half);
        const double y0 = std::max(f.y - f.h / 2, -half), y1 = std::min(f.y + f.h / 2, half);
        if (x1 > x0 && y1 > y0) boxes.push_back({x0, x1, y0, y1});
      }
      const huv_polygons& P = *rb.polys;
      for (int64_t j = P.poly_offsets[i]; j < P.poly_offsets[i + 1]; ++j) {
        std::vector<Pt> v;
        for (int64_t k = P.vertex_offsets[j]; k < P.vertex_offsets[j + 1]; ++k)
          v.push_back({P.vertices[2 * k], P.vertices[2 * k + 1]});
        if (v.size() < 3) continue;
        if (!polygon_simple(v))
          return fail(HUV_INVALID_PARAMETER, "clip " + std::to_string(i) + ", polygon " +
                                                 std::to_string(j - P.poly_offsets[i]) +
                                                 ": edges cross; rasterise it instead");
        v = clip_polygon(v, half);
        if (v.empty()) continue;
        double twice = 0.0;
        std::array<double, 4> bb{v[0].first, v[0].first, v[0].second, v[0].second};
        for (size_t k = 0; k < v.size(); ++k) {
          const Pt &a = v[k], &b = v[(k + 1) % v.size()];
          twice += a.first * b.second - b.first * a.second;
          bb = {std::min(bb[0], a.first), std::max(bb[1], a.first), std::min(bb[2], a.second),
                std::max(bb[3], a.second)};
        }
        if (twice == 0.0) continue;
        for (const auto& o : boxes)
          if (std::min(bb[1], o[1]) - std::max(bb[0], o[0]) > eps &&
              std::min(bb[3], o[3]) - std::max(bb[2], o[2]) > eps)
            return fail(HUV_INVALID_PARAMETER, "clip " + std::to_string(i) + ", polygon " +
                                                   std::to_string(j - P.poly_offsets[i]) +
                                                   ": overlaps another feature; rasterise the clip instead");
        boxes.push_back(bb);
        const huv::cd w = huv::cd(P.values[2 * j], P.values[2 * j + 1]) - bg;
        const double sg = twice > 0 ? 1.0 : -1.0;
        pdc[2 * i] += w.real() * 0.5 * std::fabs(twice);
        pdc[2 * i + 1] += w.imag() * 0.5 * std::fabs(twice);
        for (size_t k = 0; k < v.size(); ++k) {
          const Pt &a = v[k], &b = v[(k + 1) % v.size()];
          const double dx = b.first - a.first, dy = b.second - a.second;
          if (dx == 0.0 && dy == 0.0) continue;
          edges.insert(edges.end(), {dx, dy, 0.5 * (a.first + b.first), 0.5 * (a.second + b.second),
                                     sg * w.real(), sg * w.imag()});
        }
      }
    }
    eoff[i + 1] = (int64_t)(edges.size() / 6);
    if (eoff[i + 1] > kPolyMaxEdges) return fail(HUV_INVALID_PARAMETER, "too many polygon edges in one call");
  }
  const int64_t ncells = coff[nclips];
  const size_t cell_b = cells.size() * sizeof(double);
  const size_t off_b = coff.size() * sizeof(int64_t);
  const size_t bg_b = (size_t)nclips * 2 * sizeof(double);
  const int64_t nedges = eoff[nclips];
  const size_t edge_b = edges.size() * sizeof(double);
  const size_t eoff_b = nedges > 0 ? eoff.size() * sizeof(int64_t) : 0;
  const size_t pdc_b = nedges > 0 ? pdc.size() * sizeof(double) : 0;
  const size_t bytes = cell_b + off_b + bg_b + edge_b + eoff_b + pdc_b;
  hipError_t e;
  // the staging area may still feed an earlier copy on the stream
  e = hipStreamSynchronize(c->stream());
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  if (!c->staging(bytes, e)) return hip_fail(e, "staging buffer");
  if (cell_b) std::memcpy(c->pin_h, cells.data(), cell_b);
  std::memcpy(c->pin_h + cell_b, coff.data(), off_b);
  std::memcpy(c->pin_h + cell_b + off_b, rb.background, bg_b);
  char* ph = c->pin_h + cell_b + off_b + bg_b;
  if (nedges > 0) {
    std::memcpy(ph, edges.data(), edge_b);
    std::memcpy(ph + edge_b, eoff.data(), eoff_b);
    std::memcpy(ph + edge_b + eoff_b, pdc.data(), pdc_b);
  }
  e = hipMemcpyAsync(c->pin_d, c->pin_h, bytes, hipMemcpyHostToDevice, c->stream());
  if (e != hipSuccess) return hip_fail(e, "rectangle upload");
  const double* d_cells = (const double*)c->pin_d;
  const int64_t* d_coff = (const int64_t*)(c->pin_d + cell_b);
  const double* d_bg = (const double*)(c->pin_d + cell_b + off_b);
  const int nk = 2 * B + 1;
  double2* X = nullptr;
  double2* Y = nullptr;
  if (ncells > 0) {
    X = (double2*)c->ws.get(Workspace::RECT_TAB, (size_t)ncells * nk * 2 * sizeof(double2), e);
    if (!X) return hip_fail(e, "workspace");
    Y = X + (size_t)ncells * nk;
  }
  const double L = (double)n * rb.pixel_nm;
  hipEvent_t ev;
  c->stage_begin(ST_RECT_TABLES, &ev);
  e = huvk::launch_rect_tables(d_cells, ncells, n, L, B, X, Y, c->stream());
  if (e != hipSuccess) return hip_fail(e, "rect_tables launch");
  c->stage_end(ST_RECT_TABLES, ev);
  c->stage_begin(ST_RECT_BAND, &ev);
  e = huvk::launch_rect_band(X, Y, d_coff, d_bg, nclips, n, B, BB, Mq, Mp, c->stream());
  if (e != hipSuccess) return hip_fail(e, "rect_band launch");
  c->stage_end(ST_RECT_BAND, ev);
  if (nedges > 0) {
    const char* pd = c->pin_d + cell_b + off_b + bg_b;
    c->stage_begin(ST_POLY_BAND, &ev);
    e = huvk::launch_poly_band((const double*)pd, (const int64_t*)(pd + edge_b),
                               (const double*)(pd + edge_b + eoff_b), nclips, n, L, B, BB, Mq, Mp,
                               c->stream());
    if (e != hipSuccess) return hip_fail(e, "poly_band launch");
    c->stage_end(ST_POLY_BAND, ev);
  }
  return HUV_OK;
}
}  // namespace huvi

using namespace huvi;

// ===========================================================================
// Host engine
// ===========================================================================
huv_status huvi::check_build_args(const huv_optics* optics, int n, double pixel_nm,
                                  double wavelength_nm, int max_kernels,
                                  double energy_fraction) {
  if (max_kernels < 1) return fail(HUV_INVALID_PARAMETER, "max_kernels: must be at least 1");
  if (!(energy_fraction > 0.0 && energy_fraction <= 1.0))
    return fail(HUV_INVALID_PARAMETER, "kernel_energy_fraction: must be in (0, 1]");
  if (n < 2 || !(std::isfinite(pixel_nm) && pixel_nm > 0.0))
    return fail(HUV_INVALID_PARAMETER, "grid: needs size >= 2 and a positive finite pixel_nm");
  if (!(std::isfinite(wavelength_nm) && wavelength_nm > 0.0))
    return fail(HUV_INVALID_PARAMETER, "wavelength_nm: must be positive and finite");
  const double medium = optics->kind == 0 ? optics->immersion_index : 1.0;
  if (!(std::isfinite(medium) && medium >= 1.0 && optics->na > 0.0 && optics->na < medium))
    return fail(HUV_INVALID_PARAMETER,
                "numerical_aperture: must be positive and below the image-space index");
  return HUV_OK;
}

huv::Optics huvi::to_optics(const huv_optics* o) {
  huv::Optics r;
  r.kind = o->kind;
  r.na = o->na;
  r.immersion_index = o->immersion_index;
  r.flare = o->flare;
  r.central_obscuration = o->central_obscuration;
  r.transmission = o->transmission;
  r.paraxial_defocus = o->paraxial_defocus;
  return r;
}

static huv::Illumination to_ill(const huv_illumination* i) {
  huv::Illumination r;
  r.kind = i->kind;
  for (int k = 0; k < 6; ++k) r.p[k] = i->p[k];
  return r;
}

extern "C" {

const char* huv_last_error(void) { return g_err.c_str(); }
int huv_abi_version(void) { return 102; }

double huv_illumination_at(const huv_illumination* ill, double sx, double sy) {
  if (!ill) return 0.0;
  return huv::illumination_at(to_ill(ill), sx, sy);
}

huv_status huv_sample_source(const huv_illumination* ill, int points_per_axis, double df_sigma,
                             double* xyw, int64_t capacity, int64_t* count) {
  if (!ill || !count) return fail(HUV_INVALID_PARAMETER, "null argument");
  try {
    const huv::Illumination s = to_ill(ill);
    std::vector<huv::SourcePoint> pts;
    std::string err;
    const int st = huv::sample_intensity(
        [&](double x, double y) { return huv::illumination_at(s, x, y); }, points_per_axis,
        df_sigma, pts, err);
    if (st != huv::OK) return fail(st, err);
    *count = (int64_t)pts.size();
    if (xyw)
      for (int64_t i = 0; i < std::min<int64_t>(capacity, *count); ++i) {
        xyw[3 * i] = pts[i].sx;
        xyw[3 * i + 1] = pts[i].sy;
        xyw[3 * i + 2] = pts[i].weight;
      }
    return HUV_OK;
  } catch (const std::exception& ex) {
    return fail(HUV_INTERNAL_ERROR, ex.what());
  }
}

huv_status huv_pupil(const huv_optics* optics, double fx, double fy, double z, double lambda,
                     double* re, double* im) {
  if (!optics || !re || !im) return fail(HUV_INVALID_PARAMETER, "null argument");
  const huv::cd p = huv::pupil_function(to_optics(optics), fx, fy, z, lambda);
  *re = p.real();
  *im = p.imag();
  return HUV_OK;
}

int huv_select_kernel_count(const double* values, int64_t nvals, double trace,
                            int64_t max_kernels, double fraction) {
  return huv::select_kernel_count(values, nvals, trace, max_kernels, fraction);
}

huv_status huv_kernels_build(const huv_optics* optics, const double* xyw, int64_t npoints, int n,
                             double pixel_nm, double defocus_nm, double wavelength_nm,
                             int max_kernels, double energy_fraction, int method,
                             huv_host_kernels** out) {
  if (!optics || !xyw || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  const huv_status chk =
      check_build_args(optics, n, pixel_nm, wavelength_nm, max_kernels, energy_fraction);
  if (chk != HUV_OK) return chk;
  try {
    std::vector<huv::SourcePoint> pts((size_t)npoints);
    for (int64_t i = 0; i < npoints; ++i) pts[i] = {xyw[3 * i], xyw[3 * i + 1], xyw[3 * i + 2]};
    huv::BuildSettings bs;
    bs.max_kernels = max_kernels;
    bs.kernel_energy_fraction = energy_fraction;
    bs.method = method;
    auto h = std::make_unique<huv_host_kernels>();
    std::string err;
    const int st = huv::build_kernel_set(to_optics(optics), pts, n, pixel_nm, defocus_nm,
                                         wavelength_nm, bs, h->ks, err);
    if (st != huv::OK) return fail(st, err);
    *out = h.release();
    return HUV_OK;
  } catch (const std::exception& ex) {
    return fail(HUV_INTERNAL_ERROR, ex.what());
  }
}

huv_status huv_kernels_info_get(const huv_host_kernels* k, huv_kernels_info* info) {
  if (!k || !info) return fail(HUV_INVALID_PARAMETER, "null argument");
  const huv::KernelSet& s = k->ks;
  info->n = s.n;
  info->pixel_nm = s.pixel_nm;
  info->num_kernels = s.num_kernels();
  info->num_support = s.num_support();
  info->num_rows = (int64_t)s.nonzero_rows.size();
  info->method = s.method;
  info->num_columns = s.num_columns;
  info->support_exceeds_nyquist = s.support_exceeds_nyquist ? 1 : 0;
  info->captured_energy_fraction = s.captured_energy_fraction;
  info->clear_field = s.clear_field;
  info->pupil_peak = s.pupil_peak;
  info->normalization_divisor = s.normalization_divisor();
  info->defocus_nm = s.defocus_nm;
  info->wavelength_nm = s.wavelength_nm;
  return HUV_OK;
}

huv_status huv_kernels_copy(const huv_host_kernels* k, int64_t* support, int64_t* rows,
                            double* eigenvalues, double* values) {
  if (!k) return fail(HUV_INVALID_PARAMETER, "null argument");
  const huv::KernelSet& s = k->ks;
  if (support) std::copy(s.support.begin(), s.support.end(), support);
  if (rows) std::copy(s.nonzero_rows.begin(), s.nonzero_rows.end(), rows);
  if (eigenvalues) std::copy(s.eigenvalues.begin(), s.eigenvalues.end(), eigenvalues);
  if (values)
    for (size_t i = 0; i < s.values.size(); ++i) {
      values[2 * i] = s.values[i].real();
      values[2 * i + 1] = s.values[i].imag();
    }
  return HUV_OK;
}

void huv_kernels_free(huv_host_kernels* k) { delete k; }

huv_status huv_rasterize_rects(int n, double pixel_nm, int nrects, const double* rects,
                               double bg_re, double bg_im, double* out) {
  if (!out || (nrects > 0 && !rects) || n < 1 || !(pixel_nm > 0.0))
    return fail(HUV_INVALID_PARAMETER, "invalid rasterize arguments");
  std::vector<huv::RectFeature> fs((size_t)nrects);
  for (int i = 0; i < nrects; ++i) {
    const double* r = rects + 6 * i;
    fs[i] = {r[0], r[1], r[2], r[3], huv::cd(r[4], r[5])};
  }
  std::vector<huv::cd> img;
  huv::rasterize_rects(fs, huv::cd(bg_re, bg_im), n, pixel_nm, img);
  for (size_t i = 0; i < img.size(); ++i) {
    out[2 * i] = img[i].real();
    out[2 * i + 1] = img[i].imag();
  }
  return HUV_OK;
}

// ===========================================================================
// Device backend
// ===========================================================================
huv_status huv_ctx_create(int device, huv_ctx** out) {
  if (!out) return fail(HUV_INVALID_PARAMETER, "null argument");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(HUV_DEVICE_ERROR, "no HIP device available (the backend needs an MI355X GPU)");
  if (device < 0 || device >= count) return fail(HUV_INVALID_PARAMETER, "device out of range");
  e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(HUV_UNSUPPORTED, std::string("device is ") + prop.gcnArchName +
                                     "; this backend is built for gfx950 (MI355X)");
  auto c = std::make_unique<huv_ctx>();
  c->device = device;
  e = scratch_address(device, &c->fixed);
  if (e != hipSuccess) return hip_fail(e, "scratch symbol address");
  c->fixed_cap = kScratchBytes;
  // A blocking stream: it orders with work a caller puts on the null stream.
  e = hipStreamCreate(&c->own_stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
  c->active = c->own_stream;
  *out = c.release();
  return HUV_OK;
}

void huv_ctx_destroy(huv_ctx* c) { delete c; }

huv_status huv_ctx_set_stream(huv_ctx* c, void* s) {
  if (!c) return fail(HUV_INVALID_PARAMETER, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  c->active = (hipStream_t)s;
  return HUV_OK;
}

void* huv_ctx_stream(huv_ctx* c) { return c ? (void*)c->stream() : nullptr; }

void* huv_ctx_own_stream(huv_ctx* c) { return c ? (void*)c->own_stream : nullptr; }

huv_status huv_ctx_set_profiling(huv_ctx* c, int on) {
  if (!c) return fail(HUV_INVALID_PARAMETER, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  c->profiling = on != 0;
  return HUV_OK;
}

huv_status huv_ctx_stage_times(huv_ctx* c, double* ms, int64_t* calls, int nstages, int reset) {
  if (!c) return fail(HUV_INVALID_PARAMETER, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  c->collect_stage_times();
  for (int i = 0; i < nstages && i < ST_COUNT; ++i) {
    if (ms) ms[i] = c->stage_ms[i];
    if (calls) calls[i] = c->stage_calls[i];
  }
  if (reset)
    for (int i = 0; i < ST_COUNT; ++i) {
      c->stage_ms[i] = 0.0;
      c->stage_calls[i] = 0;
    }
  return HUV_OK;
}

huv_status huv_ctx_synchronize(huv_ctx* c) {
  if (!c) return fail(HUV_INVALID_PARAMETER, "null context");
  hipSetDevice(c->device);
  const hipError_t e = hipStreamSynchronize(c->stream());
  return e == hipSuccess ? HUV_OK : hip_fail(e, "hipStreamSynchronize");
}

huv_status huv_kset_upload(huv_ctx* c, int n, int K, int64_t S, const int64_t* support,
                           const double* eig, const double* values, double flare, huv_kset** out) {
  if (!c || !support || !eig || !values || !out)
    return fail(HUV_INVALID_PARAMETER, "null argument");
  if (n < 2 || (n & (n - 1)) != 0) return fail(HUV_INVALID_PARAMETER, "grid size must be a power of two");
  if (K < 1 || S < 1) return fail(HUV_INVALID_PARAMETER, "empty kernel set");
  if (!(flare >= 0.0 && flare < 1.0)) return fail(HUV_INVALID_PARAMETER, "flare must be in [0, 1)");
  int B = 0;
  for (int64_t t = 0; t < S; ++t) {
    const int64_t idx = support[t];
    if (idx < 0 || idx >= (int64_t)n * n) return fail(HUV_INVALID_PARAMETER, "support index out of range");
    const int64_t i = idx / n, j = idx % n;
    const int64_t ky = i < n / 2 ? i : i - n, kx = j < n / 2 ? j : j - n;
    B = std::max<int>(B, (int)std::max(std::llabs(ky), std::llabs(kx)));
  }
  const int m = huvk::choose_coarse(n, B);
  if (m == 0)
    return fail(HUV_UNSUPPORTED, "kernel band " + std::to_string(B) + " on a " + std::to_string(n) +
                                     " grid needs a coarse grid above 256 (or above n)");
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  static std::atomic<uint64_t> next_id{1};
  auto ks = std::make_unique<huv_kset>();
  ks->ctx = c;
  ks->id = next_id.fetch_add(1);
  ks->n = n;
  ks->K = K;
  ks->S = S;
  ks->B = B;
  ks->m = m;
  ks->BB = huvk::coarse_band(m);
  ks->flare = flare;
  const int BB = ks->BB, W = 2 * BB + 1;
  std::vector<float2> box((size_t)K * W * W, make_float2(0.f, 0.f));
  for (int k = 0; k < K; ++k)
    for (int64_t t = 0; t < S; ++t) {
      const int64_t i = support[t] / n, j = support[t] % n;
      const int64_t ky = i < n / 2 ? i : i - n, kx = j < n / 2 ? j : j - n;
      box[((size_t)k * W + (kx + BB)) * W + (ky + BB)] =
          make_float2((float)values[2 * (k * S + t)], (float)values[2 * (k * S + t) + 1]);
    }
  // Device eigenvalues carry the 1/n^4 of the two unnormalised coarse
  // transforms (|IFFT|^2 with the 1/n^2 inverse scale).
  const double n4 = std::pow((double)n, 4.0);
  std::vector<float> lam((size_t)K);
  ks->lam_host.assign(eig, eig + K);
  for (int k = 0; k < K; ++k) lam[k] = (float)(eig[k] / n4);
  ks->support.assign(support, support + S);
  ks->values.assign(values, values + 2 * K * S);
  hipError_t e = hipMalloc(&ks->kern, box.size() * sizeof(float2));
  if (e != hipSuccess) return hip_fail(e, "kernel-set allocation");
  e = hipMalloc(&ks->lam, lam.size() * sizeof(float));
  if (e != hipSuccess) return hip_fail(e, "kernel-set allocation");
  e = hipMemcpy(ks->kern, box.data(), box.size() * sizeof(float2), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "kernel-set upload");
  e = hipMemcpy(ks->lam, lam.data(), lam.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "kernel-set upload");
  *out = ks.release();
  return HUV_OK;
}

void huv_kset_destroy(huv_kset* ks) {
  if (!ks) return;
  if (ks->ctx) {
    std::lock_guard<std::mutex> lk(ks->ctx->mu);
    hipSetDevice(ks->ctx->device);
    ks->ctx->forget_kset(ks->id);
  }
  delete ks;
}

huv_status huv_kset_shape(const huv_kset* ks, int* band, int* coarse) {
  if (!ks) return fail(HUV_INVALID_PARAMETER, "null kernel set");
  if (band) *band = ks->B;
  if (coarse) *coarse = ks->m;
  return HUV_OK;
}

}  // extern "C"

namespace huvi {

huv_status validate_planes(Ctx* c, int n, int nplanes, huv_kset* const* planes) {
  if (!planes || nplanes < 1) return fail(HUV_INVALID_PARAMETER, "need at least one kernel set");
  const int B = planes[0] ? planes[0]->B : 0;
  const double flare = planes[0] ? planes[0]->flare : 0.0;
  for (int p = 0; p < nplanes; ++p) {
    const huv_kset* ks = planes[p];
    if (!ks || ks->ctx != c) return fail(HUV_INVALID_PARAMETER, "kernel set from another context");
    if (ks->n != n)
      return fail(HUV_DIMENSION_MISMATCH, "mask grid " + std::to_string(n) +
                                              " does not match the kernel-set grid " + std::to_string(ks->n));
    if (ks->B != B || ks->flare != flare)
      return fail(HUV_INVALID_PARAMETER, "all planes of one call must share band and flare");
  }
  return HUV_OK;
}

// Device copy of the per-plane (K, kernel pointer, eigenvalue pointer)
// tables, uploaded once per distinct plane list.
huv_status plane_table(Ctx* c, int nplanes, huv_kset* const* planes, const int** dK,
                       const float2* const** dkern, const float* const** dlam) {
  std::vector<uint64_t> key;
  for (int p = 0; p < nplanes; ++p) key.push_back(planes[p]->id);
  const size_t offk = 0, offp = ((size_t)nplanes * sizeof(int) + 63) / 64 * 64;
  const size_t offl = offp + (size_t)nplanes * sizeof(void*);
  const size_t tbl = offl + (size_t)nplanes * sizeof(void*);
  void* dtab = nullptr;
  auto it = c->plane_tables.find(key);
  if (it != c->plane_tables.end()) {
    dtab = it->second;
  } else {
    std::vector<char> host(tbl, 0);
    for (int p = 0; p < nplanes; ++p) {
      ((int*)(host.data() + offk))[p] = planes[p]->K;
      ((const float2**)(host.data() + offp))[p] = planes[p]->kern;
      ((const float**)(host.data() + offl))[p] = planes[p]->lam;
    }
    hipError_t e = hipMalloc(&dtab, tbl);
    if (e != hipSuccess) return hip_fail(e, "plane table");
    e = hipMemcpy(dtab, host.data(), tbl, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipFree(dtab);
      return hip_fail(e, "plane table upload");
    }
    c->plane_tables[key] = dtab;
  }
  *dK = (const int*)((char*)dtab + offk);
  *dkern = (const float2* const*)((char*)dtab + offp);
  *dlam = (const float* const*)((char*)dtab + offl);
  return HUV_OK;
}

// Kernel stacks of a plane list packed into one zero-padded buffer
// [Q][nplanes][Kmax][W][W] for coarse grid m (W = 2 BB + 1, BB = (m - 1) / 4
// >= every set's band, Q = m / 32; copy q multiplied by e^{2 pi i kx q / m} in
// FP64) with eigenvalues / n^4 in [nplanes][Kmax] (0 for padding), built from
// the host copies of the sets.
huv_status plane_pack(Ctx* c, int nplanes, huv_kset* const* planes, int m, bool paired,
                      const float2** kern, const float** lam, int* Kmax) {
  const int BB = huvk::coarse_band(m), Q = huvk::coarse_classes(m);
  std::vector<uint64_t> key;
  for (int p = 0; p < nplanes; ++p) key.push_back(planes[p]->id);
  key.push_back((uint64_t)m);
  key.push_back((uint64_t)paired);
  auto it = c->plane_packs.find(key);
  if (it == c->plane_packs.end() && c->fixed_mode)
    return fail(HUV_INVALID_PARAMETER,
                "kernel stacks of this plane list are not built yet: call huv_socs_images once "
                "with the same kernel sets before the no-allocation entry point");
  if (it == c->plane_packs.end()) {
    Ctx::PlanePack pk;
    for (int p = 0; p < nplanes; ++p) {
      pk.Kmax = std::max(pk.Kmax, planes[p]->K);
      if (planes[p]->B > BB) return fail(HUV_INTERNAL_ERROR, "kernel band exceeds the packed box");
    }
    const int W = 2 * BB + 1;
    // W x W, or column-paired (W + 1) x W: column col at [(col + 1) / 2][row][(col + 1) % 2]
    const size_t box = (size_t)(paired ? W + 1 : W) * W;
    auto at = [&](int col, int row) -> size_t {
      return paired ? ((size_t)((col + 1) >> 1) * W + row) * 2 + ((col + 1) & 1)
                    : (size_t)col * W + row;
    };
    const size_t per_class = (size_t)nplanes * pk.Kmax * box;
    std::vector<float2> hk((size_t)Q * per_class, make_float2(0.f, 0.f));
    std::vector<float> hl((size_t)nplanes * pk.Kmax, 0.f);
    for (int p = 0; p < nplanes; ++p) {
      const huv_kset* ks = planes[p];
      const int64_t n = ks->n;
      const double n4 = std::pow((double)n, 4.0);
      for (int k = 0; k < ks->K; ++k) {
        hl[(size_t)p * pk.Kmax + k] = (float)(ks->lam_host[k] / n4);
        for (int64_t t = 0; t < ks->S; ++t) {
          const int64_t i = ks->support[t] / n, j = ks->support[t] % n;
          const int64_t ky = i < n / 2 ? i : i - n, kx = j < n / 2 ? j : j - n;
          const double re = ks->values[2 * (k * ks->S + t)], im = ks->values[2 * (k * ks->S + t) + 1];
          for (int q = 0; q < Q; ++q) {
            const long e = ((kx * q) % m + m) % m;  // e^{2 pi i kx q / m}
            const double a = 2.0 * M_PI * (double)e / (double)m;
            const double cs = std::cos(a), sn = std::sin(a);
            hk[q * per_class + ((size_t)p * pk.Kmax + k) * box + at(kx + BB, ky + BB)] =
                make_float2((float)(re * cs - im * sn), (float)(re * sn + im * cs));
          }
        }
      }
    }
    hipError_t e = hipMalloc(&pk.kern, hk.size() * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(&pk.lam, hl.size() * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(pk.kern, hk.data(), hk.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(pk.lam, hl.data(), hl.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipFree(pk.kern);
      hipFree(pk.lam);
      return hip_fail(e, "plane pack upload");
    }
    it = c->plane_packs.emplace(key, pk).first;
  }
  *kern = it->second.kern;
  *lam = it->second.lam;
  *Kmax = it->second.Kmax;
  return HUV_OK;
}

// Bytes one image keeps between the stages of socs_forward: S coarse partials
// (FP32 m x m) and the intensity band ((2B + 1) x (4B + 1) complex).
size_t image_interm_bytes(int m, int B, int S) {
  return (size_t)S * m * m * sizeof(float) + (size_t)(2 * B + 1) * (4 * B + 1) * sizeof(float2);
}

// Default number of clip ranges: enough that one range's intermediates (the
// coarse images and intensity bands written by one stage and read by the
// next) stay within about 160 MB, so they are served from the device's
// last-level cache rather than HBM, but no more than keeps at least 768
// field work-groups per range (1.5 rounds of two per CU).
int clip_ranges_default(int nclips, int nplanes, int m, int B, int S) {
  const double interm = (double)nclips * nplanes * (double)image_interm_bytes(m, B, S);
  int r = (int)std::ceil(interm / 160e6);
  const int wgs = nclips * huvk::coarse_classes(m) * S;
  r = std::min(r, std::max(1, wgs / 768));
  return std::max(1, std::min(r, 16));
}

// Forward images: band spectrum -> fields -> coarse transforms -> images,
// issued per clip range (all on the context stream; huv_socs_tuning.chunks
// ranges or clip_ranges_default).  Results do not depend on the ranges: the
// kernel splits S follow the whole call (or split_batch), and every clip's
// arithmetic is the same.  Leaves the twisted band spectrum of every clip in
// Workspace::BAND (returned in *Mq_out).
huv_status socs_forward(Ctx* c, int nclips, int n, const MaskSource& src, int nplanes,
                        huv_kset* const* planes, float* out, bool pt, float2** Mq_out,
                        const huv_socs_tuning* tune, int max_splits, int split_batch,
                        const huvk::FidelityEpi* fid) {
  const int B = planes[0]->B;
  const int m = tune && tune->m ? tune->m : planes[0]->m;
  const int BB = huvk::coarse_band(m);
  const int n1 = tune ? tune->n1 : 0, xnb = tune ? tune->xnb : 0;
  const int xlanes = tune && tune->xlanes ? tune->xlanes : huvk::fields_lanes_default(m);
  if (B > BB || m > n) return fail(HUV_INVALID_PARAMETER, "coarse grid too small for the kernel band");
  if (!huvk::fields_lanes_supported(m, xlanes))
    return fail(HUV_INVALID_PARAMETER, "two x lanes per row need a coarse grid of 64 or more");
  const double flare = planes[0]->flare;
  hipError_t e;
  const int W = 2 * BB + 1;
  float2* Mq = (float2*)c->buf(Workspace::BAND, (size_t)nclips * W * W * sizeof(float2), e);
  if (!Mq) return hip_fail(e, "workspace");
  if (Mq_out) *Mq_out = Mq;
  // the two-lane field kernel reads the column-paired copy of the band
  float2* Mp = nullptr;
  if (xlanes == 2) {
    Mp = (float2*)c->buf(Workspace::BAND_PAIRED, (size_t)nclips * (W + 1) * W * sizeof(float2), e);
    if (!Mp) return hip_fail(e, "workspace");
  }
  const float2* kern;
  const float* lam;
  int Kmax;
  huv_status st = plane_pack(c, nplanes, planes, m, xlanes == 2, &kern, &lam, &Kmax);
  if (st != HUV_OK) return st;
  int S = huvk::fields_splits(m, split_batch > 0 ? split_batch : nclips, Kmax);
  if (max_splits > 0 && S > max_splits) {
    // every split must own at least one kernel (its partials are summed)
    const int ks = (Kmax + max_splits - 1) / max_splits;
    S = (Kmax + ks - 1) / ks;
  }
  // Clip ranges (one for the fidelity epilogue, whose partials are indexed by
  // image, and for the no-allocation path, whose scratch is sized for one).
  int nr = tune && tune->chunks ? tune->chunks : clip_ranges_default(nclips, nplanes, m, B, S);
  if (fid || c->fixed_mode) nr = 1;
  nr = std::max(1, std::min(nr, nclips));
  const int cc = (nclips + nr - 1) / nr;  // clips per range
  nr = (nclips + cc - 1) / cc;            // ranges actually used
  const int cimg = cc * nplanes;          // images per range
  const bool raster = src.kind == MaskSource::REAL || src.kind == MaskSource::COMPLEX;
  // Feature lists and host spectra: the band of every clip at once (their
  // stages keep no per-clip rows); rasters: per range, below.
  if (!raster) {
    st = src.kind == MaskSource::FEATURES
             ? rect_band_stage(c, nclips, n, B, BB, *src.rects, Mq, Mp)
             : spectrum_band_stage(c, nclips, n, B, BB, src.spectra, Mq, Mp);
    if (st != HUV_OK) return st;
  }
  float* coarse = (float*)c->buf(Workspace::COARSE, (size_t)S * cimg * m * m * sizeof(float), e);
  if (!coarse) return hip_fail(e, "workspace");
  const float2* tw_m = c->twiddles(m, e);
  if (!tw_m) return hip_fail(e, "twiddle table");
  const float2* tw_n = c->twiddles(n, e);
  if (!tw_n) return hip_fail(e, "twiddle table");
  const int G = 2 * B;
  float2* tmp1 = nullptr;
  const bool fused = huvk::coarse_fused_supported(m);
  if (!fused) {
    tmp1 = (float2*)c->buf(Workspace::TMP_COARSE, (size_t)cimg * (2 * G + 1) * (m / 2) * sizeof(float2), e);
    if (!tmp1) return hip_fail(e, "workspace");
  }
  float2* Ib = (float2*)c->buf(Workspace::IBAND, (size_t)cimg * (G + 1) * (2 * G + 1) * sizeof(float2), e);
  if (!Ib) return hip_fail(e, "workspace");
  const float2* box = xlanes == 2 ? Mp : Mq;
  const size_t box_c = (size_t)(xlanes == 2 ? W + 1 : W) * W;  // band box per clip
  hipEvent_t ev;
  for (int i = 0; i < nr; ++i) {
    const int c0 = i * cc, cn = std::min(cc, nclips - c0), ni = cn * nplanes;
    if (src.kind == MaskSource::REAL) {
      st = band_stage(c, cn, n, B, BB, src.raster + (size_t)c0 * n * n, Mq + (size_t)c0 * W * W,
                      pt, n1, Mp ? Mp + (size_t)c0 * (W + 1) * W : nullptr);
    } else if (src.kind == MaskSource::COMPLEX) {
      st = band_stage_complex(c, cn, n, B, BB, (const float2*)src.raster + (size_t)c0 * n * n,
                              Mq + (size_t)c0 * W * W, pt, n1,
                              Mp ? Mp + (size_t)c0 * (W + 1) * W : nullptr);
    }
    if (st != HUV_OK) return st;
    c->stage_begin(ST_FIELDS, &ev);
    e = huvk::launch_socs_fields(m, xnb, xlanes, cn, nplanes, Kmax, S, kern, lam,
                                 box + (size_t)c0 * box_c, tw_m, coarse, c->stream());
    if (e != hipSuccess) return hip_fail(e, "socs_fields launch");
    c->stage_end(ST_FIELDS, ev);
    if (fused) {
      c->stage_begin(ST_COARSE, &ev);
      e = huvk::launch_coarse_fused(m, B, ni, S, coarse, (float)flare, tw_m, Ib, c->stream());
      if (e != hipSuccess) return hip_fail(e, "coarse launch");
      c->stage_end(ST_COARSE, ev);
    } else {
      c->stage_begin(ST_COARSE_ROWS, &ev);
      e = huvk::launch_coarse_rows(m, B, ni, S, coarse, tw_m, tmp1, c->stream());
      if (e != hipSuccess) return hip_fail(e, "coarse_rows launch");
      c->stage_end(ST_COARSE_ROWS, ev);
      c->stage_begin(ST_COARSE_COLS, &ev);
      e = huvk::launch_coarse_cols(m, B, ni, (float)flare, tmp1, tw_m, Ib, c->stream());
      if (e != hipSuccess) return hip_fail(e, "coarse_cols launch");
      c->stage_end(ST_COARSE_COLS, ev);
    }
    c->stage_begin(ST_BAND_IMAGE, &ev);
    e = fid ? huvk::launch_band_image_fidelity(n, G, ni, Ib, tw_n, *fid, c->stream())
            : huvk::launch_band_image(n, G, ni, Ib, tw_n, out + (size_t)c0 * nplanes * n * n,
                                      c->stream());
    if (e != hipSuccess) return hip_fail(e, "band_image launch");
    c->stage_end(ST_BAND_IMAGE, ev);
  }
  return HUV_OK;
}

}  // namespace huvi

extern "C" {

}  // extern "C"

namespace huvi {

std::vector<int> tuning_key(int n, const huv_kset* const* planes, int nplanes) {
  int kmax = 0;
  for (int p = 0; p < nplanes; ++p) kmax = std::max(kmax, planes[p]->K);
  return {n, planes[0]->B, nplanes, kmax};
}

// Times every launch shape on this call's inputs (coordinate descent: coarse
// grid, then band-FFT factor, then x-pass block, then clip ranges) and
// records the fastest.
// The last run leaves a valid result in `out`.
huv_status autotune_socs(Ctx* c, int nclips, int n, const MaskSource& src, int nplanes,
                         huv_kset* const* planes, float* out, bool pt, huv_socs_tuning* best) {
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e != hipSuccess) return hip_fail(e, "autotune events");
  const bool prof = c->profiling;
  c->profiling = false;
  huv_status st = HUV_OK;
  auto time_one = [&](const huv_socs_tuning& t, double& ms) -> huv_status {
    double bestms = 1e30;
    for (int rep = 0; rep < 2; ++rep) {  // first run warms caches and builds tables
      hipEventRecord(a, c->stream());
      huv_status s2 = socs_forward(c, nclips, n, src, nplanes, planes, out, pt, nullptr, &t);
      if (s2 != HUV_OK) return s2;
      hipEventRecord(b, c->stream());
      hipError_t e2 = hipEventSynchronize(b);
      if (e2 != hipSuccess) return hip_fail(e2, "autotune");
      float f = 0.f;
      hipEventElapsedTime(&f, a, b);
      if (rep > 0) bestms = std::min(bestms, (double)f);
    }
    ms = bestms;
    return HUV_OK;
  };
  // explicit values throughout (0 would mean "default" in the table)
  huv_socs_tuning cur{planes[0]->m, huvk::band_fft_default(n), huvk::kFieldsXnbDefault, 1, 1};
  double cur_ms = 1e30;
  const int B = planes[0]->B;
  for (int m : {32, 64, 128, 192, 256}) {
    if (m < 4 * B + 1 || m > n) continue;
    huv_socs_tuning t{m, cur.n1, cur.xnb, cur.xlanes, cur.chunks};
    double ms;
    st = time_one(t, ms);
    if (st != HUV_OK) break;
    if (ms < cur_ms) { cur_ms = ms; cur = t; }
  }
  int n1s[4];
  const int nn1 = huvk::band_fft_options(n, n1s);
  for (int i = 0; i < nn1 && st == HUV_OK; ++i) {
    huv_socs_tuning t = cur;
    t.n1 = n1s[i];
    double ms;
    st = time_one(t, ms);
    if (st == HUV_OK && ms < cur_ms) { cur_ms = ms; cur = t; }
  }
  for (int xlanes : {1, 2}) {
    if (!huvk::fields_lanes_supported(cur.m, xlanes)) continue;
    for (int xnb : {4, 6, 8}) {
      if (st != HUV_OK) break;
      huv_socs_tuning t = cur;
      t.xlanes = xlanes;
      t.xnb = xnb;
      double ms;
      st = time_one(t, ms);
      if (st == HUV_OK && ms < cur_ms) { cur_ms = ms; cur = t; }
    }
  }
  for (int chunks : {2, 3, 4, 6, 8}) {
    if (st != HUV_OK || chunks > nclips) break;
    huv_socs_tuning t = cur;
    t.chunks = chunks;
    double ms;
    st = time_one(t, ms);
    if (st == HUV_OK && ms < cur_ms) { cur_ms = ms; cur = t; }
  }
  // leave the chosen shape's result in `out`
  if (st == HUV_OK) st = socs_forward(c, nclips, n, src, nplanes, planes, out, pt, nullptr, &cur);
  c->profiling = prof;
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (st == HUV_OK) *best = cur;
  return st;
}

// Launch shape for a call: the table entry, a fresh autotune, or defaults.
huv_status socs_shape(Ctx* c, int nclips, int n, const MaskSource& src, int nplanes,
                      huv_kset* const* planes, float* out, bool pt, huv_socs_tuning* t,
                      bool* computed) {
  *computed = false;
  const auto key = tuning_key(n, planes, nplanes);
  auto it = c->tuned.find(key);
  if (it != c->tuned.end()) {
    *t = it->second;
    return HUV_OK;
  }
  *t = huv_socs_tuning{0, 0, 0, 0, 0};
  if (!c->autotune) return HUV_OK;
  huv_status st = autotune_socs(c, nclips, n, src, nplanes, planes, out, pt, t);
  if (st != HUV_OK) return st;
  c->tuned[key] = *t;
  *computed = true;
  return HUV_OK;
}

// Runs body() eagerly, or replays / records a graph of it keyed by `key`.
template <class Body>
huv_status graph_run(Ctx* c, const std::vector<uint64_t>& key, const std::vector<uint64_t>& ksets,
                     Body body) {
  auto it = c->graphs.find(key);
  if (it != c->graphs.end() && it->second.generation == c->ws.generation) {
    hipStream_t own = c->own_stream, act = c->active;
    hipError_t e = hipSuccess;
    if (act != own) {
      hipEvent_t ev;
      e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(ev, act);
      if (e == hipSuccess) e = hipStreamWaitEvent(own, ev, 0);
      hipEventDestroy(ev);
    }
    if (e == hipSuccess) e = hipGraphLaunch(it->second.exec, own);
    if (e == hipSuccess && act != own) {
      hipEvent_t ev;
      e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(ev, own);
      if (e == hipSuccess) e = hipStreamWaitEvent(act, ev, 0);
      hipEventDestroy(ev);
    }
    return e == hipSuccess ? HUV_OK : hip_fail(e, "graph launch");
  }
  if (it != c->graphs.end()) {
    hipStreamSynchronize(c->own_stream);
    c->drop_graph(it->second);
    c->graphs.erase(it);
  }
  huv_status st = body();  // this call's result, eagerly
  if (st != HUV_OK) return st;
  // Record the same launches for later calls (capture executes nothing).
  const uint64_t gen = c->ws.generation;
  hipStream_t saved = c->active;
  const bool prof = c->profiling;
  c->active = c->own_stream;
  c->profiling = false;
  hipError_t e = hipStreamBeginCapture(c->own_stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    c->active = saved;
    c->profiling = prof;
    return hip_fail(e, "graph capture");
  }
  st = body();
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(c->own_stream, &g);
  c->active = saved;
  c->profiling = prof;
  if (st != HUV_OK || e != hipSuccess || c->ws.generation != gen) {
    if (g) hipGraphDestroy(g);
    if (st != HUV_OK) return st;
    return e != hipSuccess ? hip_fail(e, "graph capture") : fail(HUV_INTERNAL_ERROR, "workspace grew during capture");
  }
  Ctx::GraphEntry ge;
  ge.graph = g;
  ge.generation = gen;
  ge.ksets = ksets;
  e = hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    return hip_fail(e, "graph instantiate");
  }
  c->graphs[key] = ge;
  return HUV_OK;
}

struct GroupPlan {
  huv_clip_group g;
  huv_socs_tuning t;
};

huv_status socs_groups(Ctx* c, int ngroups, const huv_clip_group* groups, const huv_socs_opts* opts,
                       int tag, bool complex_masks) {
  auto source = [&](const huv_clip_group& g) {
    MaskSource m = MaskSource::real(g.masks);
    if (complex_masks) m.kind = MaskSource::COMPLEX;
    return m;
  };
  const bool pt = opts && opts->pixel_transfer;
  std::vector<GroupPlan> plan;
  std::vector<uint64_t> key{(uint64_t)tag, (uint64_t)pt, (uint64_t)ngroups};
  std::vector<uint64_t> ksets;
  for (int i = 0; i < ngroups; ++i) {
    const huv_clip_group& g = groups[i];
    if (g.nclips == 0) continue;
    GroupPlan gp{g, {0, 0, 0, 0}};
    bool computed = false;
    huv_status st = socs_shape(c, g.nclips, g.n, source(g), g.nplanes, g.planes, g.out, pt,
                               &gp.t, &computed);
    if (st != HUV_OK) return st;
    plan.push_back(gp);
    key.insert(key.end(), {(uint64_t)g.n, (uint64_t)g.nclips, (uint64_t)(uintptr_t)g.masks,
                           (uint64_t)(uintptr_t)g.out, (uint64_t)g.nplanes, (uint64_t)gp.t.m,
                           (uint64_t)gp.t.n1, (uint64_t)gp.t.xnb, (uint64_t)gp.t.xlanes,
                           (uint64_t)gp.t.chunks});
    for (int p = 0; p < g.nplanes; ++p) {
      key.push_back(g.planes[p]->id);
      ksets.push_back(g.planes[p]->id);
    }
  }
  auto body = [&]() -> huv_status {
    for (const auto& gp : plan) {
      huv_status st = socs_forward(c, gp.g.nclips, gp.g.n, source(gp.g), gp.g.nplanes,
                                   gp.g.planes, gp.g.out, pt, nullptr, &gp.t);
      if (st != HUV_OK) return st;
    }
    return HUV_OK;
  };
  if (opts && opts->graph) return graph_run(c, key, ksets, body);
  return body();
}

}  // namespace huvi

extern "C" {

huv_status huv_socs_images(huv_ctx* c, int nclips, int n, const float* masks, int nplanes,
                           huv_kset* const* planes, float* out, const huv_socs_opts* opts) {
  if (!c || !masks || !planes || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (nclips < 0 || nplanes < 1) return fail(HUV_INVALID_PARAMETER, "need nclips >= 0 and nplanes >= 1");
  huv_status st = validate_planes(c, n, nplanes, planes);
  if (st != HUV_OK) return st;
  if (nclips == 0) return HUV_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  huv_clip_group g{n, nclips, masks, nplanes, planes, out};
  return socs_groups(c, 1, &g, opts, 1, false);
}

huv_status huv_socs_images_complex(huv_ctx* c, int nclips, int n, const float* masks,
                                   int nplanes, huv_kset* const* planes, float* out,
                                   const huv_socs_opts* opts) {
  if (!c || !masks || !planes || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (nclips < 0 || nplanes < 1) return fail(HUV_INVALID_PARAMETER, "need nclips >= 0 and nplanes >= 1");
  huv_status st = validate_planes(c, n, nplanes, planes);
  if (st != HUV_OK) return st;
  if (nclips == 0) return HUV_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  huv_clip_group g{n, nclips, masks, nplanes, planes, out};
  return socs_groups(c, 1, &g, opts, 3, true);
}

huv_status huv_socs_images_spectrum(huv_ctx* c, int nclips, int n, const double* spectra,
                                    int nplanes, huv_kset* const* planes, float* out) {
  if (!c || !spectra || !planes || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (nclips < 0 || nplanes < 1) return fail(HUV_INVALID_PARAMETER, "need nclips >= 0 and nplanes >= 1");
  huv_status st = validate_planes(c, n, nplanes, planes);
  if (st != HUV_OK) return st;
  if (nclips == 0) return HUV_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  auto it = c->tuned.find(tuning_key(n, planes, nplanes));
  MaskSource src;
  src.kind = MaskSource::SPECTRUM;
  src.spectra = spectra;
  return socs_forward(c, nclips, n, src, nplanes, planes, out, false, nullptr,
                      it != c->tuned.end() ? &it->second : nullptr);
}

huv_status huv_socs_images_noalloc(huv_ctx* c, int nclips, int n, const float* masks,
                                   int nplanes, huv_kset* const* planes, float* out,
                                   int pixel_transfer) {
  if (!c || !masks || !planes || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (nclips < 0 || nplanes < 1) return fail(HUV_INVALID_PARAMETER, "need nclips >= 0 and nplanes >= 1");
  huv_status st = validate_planes(c, n, nplanes, planes);
  if (st != HUV_OK) return st;
  if (nclips == 0) return HUV_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  if (!c->fixed) return fail(HUV_INTERNAL_ERROR, "no scratch array on this device");
  const bool pt = pixel_transfer != 0;
  // Launch shape: the context's table entry for this shape class, else the
  // defaults (no autotuning here: it times extra launches).
  huv_socs_tuning t{0, 0, 0, 0};
  auto tit = c->tuned.find(tuning_key(n, planes, nplanes));
  if (tit != c->tuned.end()) t = tit->second;
  const int B = planes[0]->B;
  const int m = t.m ? t.m : planes[0]->m;
  const int xlanes = t.xlanes ? t.xlanes : huvk::fields_lanes_default(m);
  if (!c->tw.count(n) || !c->tw.count(m) || !c->sinc.count(sinc_key(n, B, pt)))
    return fail(HUV_INVALID_PARAMETER,
                "transform tables for this grid are not built yet: call huv_socs_images once "
                "with the same kernel sets before the no-allocation entry point");
  int Kmax = 0;
  for (int p = 0; p < nplanes; ++p) Kmax = std::max(Kmax, planes[p]->K);
  // As many kernel splits as huv_socs_images would use, fewer if the scratch
  // array is too small for their partial images.
  int S = huvk::fields_splits(m, nclips, Kmax);
  size_t need = 0;
  for (;; --S) {
    const int ks = (Kmax + S - 1) / S, s_eff = (Kmax + ks - 1) / ks;
    need = socs_sizes(nclips, n, B, m, xlanes == 2, nplanes, s_eff, true).total();
    if (need <= c->fixed_cap || S == 1) break;
  }
  if (need > c->fixed_cap)
    return fail(HUV_INVALID_PARAMETER,
                "this batch needs " + std::to_string((need + (1 << 20) - 1) >> 20) +
                    " MiB of scratch; the no-allocation entry point has " +
                    std::to_string(c->fixed_cap >> 20) + " MiB per GPU: pass fewer clips per call");
  const bool prof = c->profiling;
  c->profiling = false;  // per-stage timing creates events
  c->fixed_mode = true;
  c->fixed_used = 0;
  st = socs_forward(c, nclips, n, MaskSource::real(masks), nplanes, planes, out, pt, nullptr, &t, S);
  c->fixed_mode = false;
  c->profiling = prof;
  return st;
}

huv_status huv_socs_images_features(huv_ctx* c, int nclips, int n, double pixel_nm,
                                    const int64_t* offsets, const double* rects,
                                    const huv_polygons* polys, const double* background,
                                    int nplanes, huv_kset* const* planes, float* out,
                                    const huv_socs_opts* opts) {
  (void)opts;  // the spectrum is exact (no pixel transfer); no graph capture
  if (!c || !planes || !out || !offsets || !background)
    return fail(HUV_INVALID_PARAMETER, "null argument");
  if (nclips < 0 || nplanes < 1) return fail(HUV_INVALID_PARAMETER, "need nclips >= 0 and nplanes >= 1");
  huv_status st = validate_planes(c, n, nplanes, planes);
  if (st != HUV_OK) return st;
  RectBatch rb{pixel_nm, offsets, rects, background};
  rb.polys = polys;
  st = validate_rects(nclips, n, rb);
  if (st != HUV_OK) return st;
  if (nclips == 0) return HUV_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  // a launch shape tuned for these planes applies (its band-FFT factor is unused)
  auto it = c->tuned.find(tuning_key(n, planes, nplanes));
  MaskSource src;
  src.kind = MaskSource::FEATURES;
  src.rects = &rb;
  return socs_forward(c, nclips, n, src, nplanes, planes, out, false, nullptr,
                      it != c->tuned.end() ? &it->second : nullptr);
}

huv_status huv_socs_images_rects(huv_ctx* c, int nclips, int n, double pixel_nm,
                                 const int64_t* offsets, const double* rects,
                                 const double* background, int nplanes, huv_kset* const* planes,
                                 float* out, const huv_socs_opts* opts) {
  return huv_socs_images_features(c, nclips, n, pixel_nm, offsets, rects, nullptr, background,
                                  nplanes, planes, out, opts);
}

huv_status huv_socs_images_batch(huv_ctx* c, int ngroups, const huv_clip_group* groups,
                                 const huv_socs_opts* opts) {
  if (!c || (ngroups > 0 && !groups)) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (ngroups < 0) return fail(HUV_INVALID_PARAMETER, "need ngroups >= 0");
  for (int i = 0; i < ngroups; ++i) {
    const huv_clip_group& g = groups[i];
    if (g.nclips < 0 || g.nplanes < 1) return fail(HUV_INVALID_PARAMETER, "group: need nclips >= 0 and nplanes >= 1");
    if (g.nclips > 0 && (!g.masks || !g.out || !g.planes)) return fail(HUV_INVALID_PARAMETER, "group: null buffer");
    if (g.nclips == 0) continue;
    huv_status st = validate_planes(c, g.n, g.nplanes, g.planes);
    if (st != HUV_OK) return st;
  }
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  return socs_groups(c, ngroups, groups, opts, 2, false);
}

huv_status huv_ctx_set_autotune(huv_ctx* c, int on) {
  if (!c) return fail(HUV_INVALID_PARAMETER, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  c->autotune = on != 0;
  return HUV_OK;
}

huv_status huv_ctx_tuning_get(huv_ctx* c, int n, int band, int nplanes, int kmax,
                              huv_socs_tuning* out, int* found) {
  if (!c || !out || !found) return fail(HUV_INVALID_PARAMETER, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  auto it = c->tuned.find({n, band, nplanes, kmax});
  *found = it != c->tuned.end();
  *out = *found ? it->second : huv_socs_tuning{0, 0, 0, 0, 0};
  return HUV_OK;
}

huv_status huv_ctx_tuning_set(huv_ctx* c, int n, int band, int nplanes, int kmax,
                              const huv_socs_tuning* t) {
  if (!c || !t) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (t->m != 0 && (t->m < 4 * band + 1 || t->m > n ||
                    (t->m != 32 && t->m != 64 && t->m != 128 && t->m != 192 && t->m != 256)))
    return fail(HUV_INVALID_PARAMETER,
                "tuning: coarse grid must be one of 32, 64, 128, 192, 256, >= 4 band + 1 and <= n");
  if (t->n1 != 0) {
    int n1s[4];
    const int k = huvk::band_fft_options(n, n1s);
    if (std::find(n1s, n1s + k, t->n1) == n1s + k)
      return fail(HUV_INVALID_PARAMETER, "tuning: unsupported band-FFT factor for this n");
  }
  if (t->xnb != 0 && t->xnb != 4 && t->xnb != 6 && t->xnb != 8)
    return fail(HUV_INVALID_PARAMETER, "tuning: xnb must be 4, 6 or 8");
  if (t->xlanes < 0 || t->xlanes > 2 ||
      (t->xlanes == 2 && t->m != 0 && !huvk::fields_lanes_supported(t->m, 2)))
    return fail(HUV_INVALID_PARAMETER, "tuning: xlanes must be 1, or 2 with a coarse grid of 64 or more");
  if (t->chunks < 0 || t->chunks > 16)
    return fail(HUV_INVALID_PARAMETER, "tuning: chunks must be 0 (default) or 1 to 16");
  std::lock_guard<std::mutex> lk(c->mu);
  c->tuned[{n, band, nplanes, kmax}] = *t;
  return HUV_OK;
}

huv_status huv_socs_images_host(huv_ctx* c, int nclips, int n, const float* masks, int nplanes,
                                huv_kset* const* planes, float* out, const huv_socs_opts* opts) {
  if (!c || !masks || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  hipSetDevice(c->device);
  const size_t in_b = (size_t)nclips * n * n * sizeof(float);
  const size_t out_b = (size_t)nclips * nplanes * n * n * sizeof(float);
  float *dm = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&dm, in_b);
  if (e != hipSuccess) return hip_fail(e, "input allocation");
  e = hipMalloc(&dout, out_b);
  if (e != hipSuccess) {
    hipFree(dm);
    return hip_fail(e, "output allocation");
  }
  huv_status st = HUV_OK;
  e = hipMemcpy(dm, masks, in_b, hipMemcpyHostToDevice);
  if (e != hipSuccess) st = hip_fail(e, "input copy");
  // Per-call device buffers: never recorded into a graph.
  huv_socs_opts o{opts ? opts->pixel_transfer : 0, 0};
  if (st == HUV_OK) st = huv_socs_images(c, nclips, n, dm, nplanes, planes, dout, &o);
  if (st == HUV_OK) {
    e = hipStreamSynchronize(c->stream());
    if (e == hipSuccess) e = hipMemcpy(out, dout, out_b, hipMemcpyDeviceToHost);
    if (e != hipSuccess) st = hip_fail(e, "output copy");
  }
  hipFree(dm);
  hipFree(dout);
  return st;
}

huv_status huv_band_spectrum(huv_ctx* c, int nclips, int n, int band, const float* masks,
                             float* out, int pixel_transfer) {
  if (!c || !masks || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (n < 256 || n > 2048 || (n & (n - 1)) != 0)
    return fail(HUV_UNSUPPORTED, "band spectrum supports n in {256, 512, 1024, 2048}");
  if (band < 0 || 2 * band + 1 > n) return fail(HUV_INVALID_PARAMETER, "band out of range");
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  return band_stage(c, nclips, n, band, band, masks, (float2*)out, pixel_transfer != 0);
}

huv_status huv_band_spectrum_features(huv_ctx* c, int nclips, int n, double pixel_nm, int band,
                                      const int64_t* offsets, const double* rects,
                                      const huv_polygons* polys, const double* background,
                                      float* out) {
  if (!c || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (nclips < 0) return fail(HUV_INVALID_PARAMETER, "need nclips >= 0");
  if (band < 0 || 2 * band + 1 > n) return fail(HUV_INVALID_PARAMETER, "band out of range");
  if (polys && band > huvk::kPolyMaxBand)
    return fail(HUV_UNSUPPORTED, "polygon masks support band half-widths up to " +
                                     std::to_string(huvk::kPolyMaxBand));
  RectBatch rb{pixel_nm, offsets, rects, background};
  rb.polys = polys;
  huv_status st = validate_rects(nclips, n, rb);
  if (st != HUV_OK) return st;
  if (nclips == 0) return HUV_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  return rect_band_stage(c, nclips, n, band, band, rb, (float2*)out, nullptr);
}

huv_status huv_band_spectrum_rects(huv_ctx* c, int nclips, int n, double pixel_nm, int band,
                                   const int64_t* offsets, const double* rects,
                                   const double* background, float* out) {
  return huv_band_spectrum_features(c, nclips, n, pixel_nm, band, offsets, rects, nullptr,
                                    background, out);
}

}  // extern "C"
