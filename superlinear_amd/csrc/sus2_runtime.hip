// Model upload, workspace management and the neighbour-count scan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <string>

#include "../../include/sus2amd.h"
#include "sus2_runtime.hpp"

namespace sus2 {

constexpr int REPORT_WORDS = 8;  // 64-bit words of the coherent host report

#define RT_CHECK(x)                                                                                \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) throw Error(SUS2_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {
thread_local Arena* t_arena = nullptr;
thread_local bool t_capturing = false;
}
Arena* current_arena() { return t_arena; }
bool capturing() { return t_capturing; }
CaptureScope::CaptureScope() { t_capturing = true; }
CaptureScope::~CaptureScope() { t_capturing = false; }
ArenaScope::ArenaScope(Arena* a) : prev(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev; }

DevBuf::~DevBuf() {
  if (p && owned) (void)hipFree(p);
}

void DevBuf::ensure(size_t n) {
  if (n <= bytes) return;
  if (Arena* a = t_arena) {  // carve (256-byte aligned) from the arena; nothing is freed
    const size_t want = ((std::max<size_t>(n, 1) + 255) / 256) * 256;
    if (want > a->cap - a->used)
      throw Error(SUS2_E_RANGE, "batch needs more device scratch than the " + std::to_string(a->cap >> 20) +
                                    " MiB static region of the allocation-free entry point (" +
                                    std::to_string((a->used + want) >> 20) + " MiB requested so far)");
    if (p && owned) RT_CHECK(hipFree(p));
    p = a->base + a->used;
    a->used += want;
    bytes = want;
    owned = false;
    return;
  }
  if (t_capturing) throw Error(SUS2_E_INVALID, "device buffer growth during HIP graph capture");
  if (p && owned) RT_CHECK(hipFree(p));
  p = nullptr;
  bytes = 0;
  owned = true;
  const size_t want = std::max<size_t>(n + n / 4, 256);
  RT_CHECK(hipMalloc(&p, want));
  bytes = want;
}

Profiler::~Profiler() {
  for (auto& p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto e : pool) (void)hipEventDestroy(e);
}

int Profiler::begin(const std::string& name) {
  if (!on || t_capturing) return -1;
  int id = -1;
  for (size_t k = 0; k < names.size(); ++k)
    if (names[k] == name) { id = (int)k; break; }
  if (id < 0) {
    id = (int)names.size();
    names.push_back(name);
    total_ms.push_back(0);
    launches.push_back(0);
  }
  hipEvent_t ev[2];
  for (auto& e : ev) {
    if (!pool.empty()) { e = pool.back(); pool.pop_back(); }
    else RT_CHECK(hipEventCreate(&e));
  }
  pending.push_back({id, ev[0], ev[1]});
  return (int)pending.size() - 1;
}

void Profiler::collect_ready() {
  if (pending.size() < 4096) return;  // bound the number of outstanding events
  collect_all();
}

void Profiler::collect_all() {
  for (auto& p : pending) {
    RT_CHECK(hipEventSynchronize(p.b));
    float ms = 0;
    RT_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    total_ms[p.id] += ms;
    launches[p.id] += 1;
    pool.push_back(p.a);
    pool.push_back(p.b);
  }
  pending.clear();
}

void Profiler::reset() {
  collect_all();
  names.clear();
  total_ms.clear();
  launches.clear();
}

Workspace::~Workspace() {
  if (host_total) (void)hipHostFree(host_total);
}

void Workspace::ensure_atoms(int n) {
  nbr_cnt.ensure(sizeof(int) * (size_t)std::max(n, 1));
  nbr_off.ensure(sizeof(int) * ((size_t)n + 1));
  if (t_arena) return;  // allocation-free calls read results by plain copies instead of pinned pages
  if (!host_total) RT_CHECK(hipHostMalloc((void**)&host_total, 4 * sizeof(int64_t), hipHostMallocDefault));
  if (!host_report) {
    RT_CHECK(hipHostMalloc((void**)&host_report, REPORT_WORDS * sizeof(int64_t), hipHostMallocCoherent));
    for (int k = 0; k < REPORT_WORDS; ++k) host_report[k] = 0;
  }
}

// Single-workgroup exclusive scan over tiles of SCAN_TILE counts: each thread
// loads 16 consecutive counts (four 16-byte loads), a wave
// prefix sum and a scan of the 16 wave sums give the tile's offsets, and the
// running total carries to the next tile.  Totals above 2^31 are reported
// through the int64 total.
constexpr int SCAN_THREADS = 1024, SCAN_PER = 16, SCAN_TILE = SCAN_THREADS * SCAN_PER;
static_assert(SCAN_PER % 4 == 0, "whole int4 groups per thread");
constexpr int SCAN_WAVES = SCAN_THREADS / 64;

__global__ __launch_bounds__(SCAN_THREADS) void k_scan(int n, int* __restrict__ cnt, int* __restrict__ off,
                                                       int64_t* __restrict__ total, int cap_lim, long edge_lim,
                                                       int64_t* report, int64_t seq, int* sticky) {
  __shared__ long wsum[SCAN_WAVES];
  __shared__ int wmax[SCAN_WAVES];
  const int t = threadIdx.x, lane = t % 64, w = t / 64;
  // 16-byte loads and stores where both arrays are 16-byte aligned (i0 is a
  // multiple of 16 elements)
  const bool vec = ((reinterpret_cast<uintptr_t>(cnt) | reinterpret_cast<uintptr_t>(off)) & 15) == 0;
  long carry = 0;
  int mx = 0;
  for (int base = 0; base < n; base += SCAN_TILE) {
    const int i0 = base + t * SCAN_PER;
    int v[SCAN_PER];
    const bool whole = vec && i0 + SCAN_PER <= n;
    if (whole) {
#pragma unroll
      for (int q = 0; q < SCAN_PER / 4; ++q) {
        const int4 x = reinterpret_cast<const int4*>(cnt + i0)[q];
        v[4 * q] = x.x;
        v[4 * q + 1] = x.y;
        v[4 * q + 2] = x.z;
        v[4 * q + 3] = x.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < SCAN_PER; ++k) v[k] = i0 + k < n ? cnt[i0 + k] : 0;
    }
    long s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
      s += v[k];
      mx = max(mx, v[k]);
    }
    long incl = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const long y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    if (w == 0) {
      long x = lane < SCAN_WAVES ? wsum[lane] : 0;
#pragma unroll
      for (int d = 1; d < SCAN_WAVES; d <<= 1) {
        const long y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
      }
      if (lane < SCAN_WAVES) wsum[lane] = x;
    }
    __syncthreads();
    long run = carry + (w ? wsum[w - 1] : 0) + incl - s;
    if (whole) {
#pragma unroll
      for (int q = 0; q < SCAN_PER / 4; ++q) {
        int4 y;
        y.x = (int)run;
        run += v[4 * q];
        y.y = (int)run;
        run += v[4 * q + 1];
        y.z = (int)run;
        run += v[4 * q + 2];
        y.w = (int)run;
        run += v[4 * q + 3];
        reinterpret_cast<int4*>(off + i0)[q] = y;
      }
    } else {
#pragma unroll
      for (int k = 0; k < SCAN_PER; ++k)
        if (i0 + k < n) {
          off[i0 + k] = (int)run;
          run += v[k];
        }
    }
    carry += wsum[SCAN_WAVES - 1];
    __syncthreads();  // wsum is rewritten by the next tile
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) mx = max(mx, __shfl_xor(mx, d, 64));
  if (lane == 0) wmax[w] = mx;
  __syncthreads();
  int m = 0;
#pragma unroll
  for (int k = 0; k < SCAN_WAVES; ++k) m = max(m, wmax[k]);
  const bool over = carry > edge_lim || m > cap_lim;
  if (over)
    for (int i = t; i < n; i += SCAN_THREADS) {
      off[i] = 0;
      cnt[i] = 0;
    }
  if (t == 0) {
    off[n] = over ? 0 : (int)carry;
    total[0] = carry;
    total[1] = m;
    total[2] = over;
    if (sticky && over) *sticky = 1;
    if (report) {  // two self-tagged words, each one atomic 64-bit store at system scope (no fence)
      const uint64_t tag = (uint64_t)(seq & 0x7fffffff) << 32;
      __hip_atomic_store(report, (int64_t)(tag | (uint64_t)(uint32_t)m), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(report + 1,
                         (int64_t)(tag | ((uint64_t)over << 31) | (uint64_t)(uint32_t)min(carry, (long)0x7fffffff)),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

void device_scan(int* cnt, int* off, int n, int64_t* total, hipStream_t st, int cap_lim, long edge_lim,
                 int64_t* report, int64_t seq, int* sticky) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_THREADS), 0, st, n, cnt, off, total, cap_lim, edge_lim, report, seq,
                     sticky);
  RT_CHECK(hipGetLastError());
}

int64_t* report_word(Workspace& ws, int k) {
  if (!ws.host_report) ws.ensure_atoms(0);
  int64_t* report = nullptr;
  RT_CHECK(hipHostGetDevicePointer((void**)&report, ws.host_report, 0));
  return report + k;
}

// How long the host spins on a report word before it waits for the stream:
// longer than a queued call of ordinary size, so that a host one call ahead
// of the device does not drain the stream (which would leave the device idle
// while the next call is launched).
constexpr auto SPIN_LIMIT = std::chrono::milliseconds(50);

void poll_report(Workspace& ws, int k, int n, int64_t seq, hipStream_t st, const char* what, uint64_t* w) {
  if (t_capturing || ws.capture) throw Error(SUS2_E_INVALID, std::string(what) + ": host wait in capture-safe mode");
  volatile int64_t* rep = ws.host_report + k;
  const uint64_t want = (uint64_t)(seq & 0x7fffffff);
  auto arrived = [&] {
    bool ok = true;
    for (int q = 0; q < n; ++q) {
      w[q] = (uint64_t)rep[q];
      ok = ok && (w[q] >> 32) == want;
    }
    return ok;
  };
  const auto t0 = std::chrono::steady_clock::now();
  bool waited = false;
  for (long polls = 1; !arrived(); ++polls) {
    if (polls % 256) continue;
    if (!waited && std::chrono::steady_clock::now() - t0 > SPIN_LIMIT) {
      waited = true;  // a long queue: wait for the stream instead of spinning
      const hipError_t e = hipStreamSynchronize(st);
      if (e != hipSuccess) throw Error(SUS2_E_DEVICE, std::string(what) + " failed: " + hipGetErrorString(e));
      continue;
    }
    if (polls % 1024) continue;
    const hipError_t e = hipStreamQuery(st);
    if (e == hipErrorNotReady) continue;
    if (e != hipSuccess) throw Error(SUS2_E_DEVICE, std::string(what) + " failed: " + hipGetErrorString(e));
    std::atomic_thread_fence(std::memory_order_seq_cst);
    if (!arrived()) throw Error(SUS2_E_DEVICE, std::string(what) + ": report missing after the stream drained");
  }
}

long exclusive_scan(Workspace& ws, int n, hipStream_t st, long* maxcnt) {
  if (t_capturing || ws.capture) throw Error(SUS2_E_INVALID, "neighbour sizing needs a host wait in capture-safe mode");
  ws.scan_tmp.ensure(8 * sizeof(int64_t));
  device_scan(ws.nbr_cnt.get<int>(), ws.nbr_off.get<int>(), n, ws.scan_tmp.get<int64_t>(), st);
  int64_t local[2];
  int64_t* dst = ws.host_total ? ws.host_total : local;
  RT_CHECK(hipMemcpyAsync(dst, ws.scan_tmp.get<int64_t>(), 2 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RT_CHECK(hipStreamSynchronize(st));
  if (maxcnt) *maxcnt = (long)dst[1];
  return (long)dst[0];
}

namespace {
struct IntPack {
  std::vector<int> data;
  size_t add(const std::vector<int>& v) {
    const size_t o = data.size();
    data.insert(data.end(), v.begin(), v.end());
    if (v.empty()) data.push_back(0);
    return o;
  }
};

template <class T>
void fill_view(DevModelT<T>& dv, const Model& m, const int* ibase, const std::vector<size_t>& io) {
  // coefficient arrays in T
  std::vector<T> th(m.theta.begin(), m.theta.end());
  dv.theta.ensure(sizeof(T) * th.size());
  RT_CHECK(hipMemcpy(dv.theta.p, th.data(), sizeof(T) * th.size(), hipMemcpyHostToDevice));
  // product-graph rows as 16-byte records: gate forward, outer forward, gate
  // reverse, outer reverse
  std::vector<GRow<T>> rows;
  size_t roff[4];
  for (int part = 0; part < 4; ++part) {
    const PassPlan& pp = part % 2 == 0 ? m.gate_plan : m.outer_plan;
    roff[part] = rows.size();
    if (part < 2)
      for (size_t r = 0; r < pp.fw_a.size(); ++r) rows.push_back(GRow<T>{pp.fw_a[r], pp.fw_b[r], T(pp.fw_c[r])});
    else
      for (size_t r = 0; r < pp.rv_out.size(); ++r)
        rows.push_back(GRow<T>{pp.rv_out[r], pp.rv_partner[r], T(pp.rv_c[r])});
  }
  rows.push_back(GRow<T>{0, 0, T(0)});
  dv.rows.ensure(sizeof(GRow<T>) * rows.size());
  RT_CHECK(hipMemcpy(dv.rows.p, rows.data(), sizeof(GRow<T>) * rows.size(), hipMemcpyHostToDevice));
  DModel<T>& v = dv.view;
  v.C = m.C; v.L = m.L; v.H = m.H; v.K = m.K; v.F = m.F; v.B = m.B; v.M = m.M; v.Q = m.Q; v.Kw = m.Kw;
  v.smode = m.smode; v.R = m.R; v.O = m.O; v.Qg = m.Qg;
  v.rc = T(m.rc); v.rca = T(m.rca); v.scaling = T(m.scaling);
  v.gated = m.gated; v.species_readout = m.species_readout;
  v.scalar_unique = 1;
  for (int n : m.scalar_map)
    if (std::count(m.scalar_map.begin(), m.scalar_map.end(), n) != 1) v.scalar_unique = 0;
  v.theta = dv.theta.template get<T>();
  v.o_shift = m.o_shift; v.o_scal = m.o_scal; v.o_rad = m.o_rad; v.o_grad = m.o_grad; v.o_gscal = m.o_gscal;
  v.o_alpha = m.o_alpha; v.o_gtype = m.o_gtype; v.o_W = m.o_W; v.o_gw = m.o_gw; v.o_ow = m.o_ow;
  v.o_species = m.o_species; v.o_read = m.o_read;
  size_t k = 0;
  v.basic_node = ibase + io[k++];
  v.scalar_map = ibase + io[k++];
  v.gate_nodes = ibase + io[k++];
  for (int pass = 0; pass < 2; ++pass) {
    const PassPlan& pp = pass == 0 ? m.gate_plan : m.outer_plan;
    DPass<T>& dp = pass == 0 ? v.gp : v.op;
    dp.nfl = (int)pp.fw_layer_off.size() - 1;
    dp.nrl = (int)pp.rv_layer_off.size() - 1;
    if (dp.nfl < 0) dp.nfl = 0;
    if (dp.nrl < 0) dp.nrl = 0;
    dp.rank = pp.rank;
    dp.begin = pp.begin;
    dp.node_lo = pp.node_lo;
    dp.max_fw_rows = pp.max_fw_rows;
    dp.max_rv_rows = pp.max_rv_rows;
    dp.node_cnt = pp.node_cnt;
    dp.nfout = pp.fw_out.empty() ? 0 : (int)pp.fw_out.size();
    dp.nrnode = pp.rv_node.empty() ? 0 : (int)pp.rv_node.size();
    dp.fl_off = ibase + io[k++]; dp.f_out = ibase + io[k++]; dp.f_roff = ibase + io[k++];
    dp.rl_off = ibase + io[k++]; dp.r_node = ibase + io[k++]; dp.r_roff = ibase + io[k++];
    dp.fl_bnd = reinterpret_cast<const int4*>(ibase + io[k++]);
    dp.rl_bnd = reinterpret_cast<const int4*>(ibase + io[k++]);
    dp.f_row = dv.rows.template get<GRow<T>>() + roff[pass];
    dp.r_row = dv.rows.template get<GRow<T>>() + roff[2 + pass];
  }
}
}  // namespace

void DevModelSet::upload(const Model& m) {
  host = &m;
  for (int n : m.gate_nodes)
    if (std::count(m.gate_nodes.begin(), m.gate_nodes.end(), n) != 1)
      throw Error(SUS2_E_UNSUPPORTED, "SUS2 model: repeated gate scalar node");
  IntPack pk;
  std::vector<size_t> io;
  io.push_back(pk.add(m.basic_node));
  io.push_back(pk.add(m.scalar_map));
  io.push_back(pk.add(m.gate_nodes));
  for (int pass = 0; pass < 2; ++pass) {
    const PassPlan& pp = pass == 0 ? m.gate_plan : m.outer_plan;
    for (const auto* vec : {&pp.fw_layer_off, &pp.fw_out, &pp.fw_row_off, &pp.rv_layer_off, &pp.rv_node,
                            &pp.rv_row_off})
      io.push_back(pk.add(*vec));
    // layer bounds, 16-byte aligned in the packed array
    auto bounds = [](const std::vector<int>& lay, const std::vector<int>& roff) {
      std::vector<int> b;
      for (size_t l = 0; l + 1 < lay.size(); ++l) {
        const int a0 = lay[l], a1 = lay[l + 1];
        b.insert(b.end(), {a0, a1, roff.empty() ? 0 : roff[a0], roff.empty() ? 0 : roff[a1]});
      }
      return b;
    };
    for (const auto& b : {bounds(pp.fw_layer_off, pp.fw_row_off), bounds(pp.rv_layer_off, pp.rv_row_off)}) {
      while (pk.data.size() % 4) pk.data.push_back(0);
      io.push_back(pk.add(b));
    }
  }
  ints.ensure(sizeof(int) * pk.data.size());
  RT_CHECK(hipMemcpy(ints.p, pk.data.data(), sizeof(int) * pk.data.size(), hipMemcpyHostToDevice));
  fill_view(d, m, ints.get<int>(), io);
  fill_view(f, m, ints.get<int>(), io);
}

void DevModelSet::upload_theta(const Model& m) {
  // kernels queued by earlier calls (on any stream) may still read them
  RT_CHECK(hipDeviceSynchronize());
  std::vector<double> td(m.theta.begin(), m.theta.end());
  std::vector<float> tf(m.theta.begin(), m.theta.end());
  RT_CHECK(hipMemcpy(d.theta.p, td.data(), sizeof(double) * td.size(), hipMemcpyHostToDevice));
  RT_CHECK(hipMemcpy(f.theta.p, tf.data(), sizeof(float) * tf.size(), hipMemcpyHostToDevice));
}

}  // namespace sus2
