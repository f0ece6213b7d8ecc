// solver.hip -- gfx950 kernels of the FIM activation-time solve and the
// device half of the C ABI (fim_create ... fim_destroy).
//
// Two modes share one pack (host.hpp):
//
//  converge (FP32, exact local solver): per global iteration a persistent
//    kernel (k_conv_band) works through the (instance, submesh) pairs of the
//    iteration's work list, dealt out per XCD in units of a few instances of
//    one submesh.  A work item stages the submesh's vertex times in LDS and
//    runs Jacobi sweeps (capped per activation): sources mark their incident
//    elements through a padded vertex -> element adjacency, the marked
//    elements evaluate exact candidates from 48-byte element records read
//    from L2, and candidates merge with LDS atomicMin on the float bit
//    pattern (times are >= 0, so the IEEE order equals the unsigned order).
//    A work item writes its lowered interface copies as (entry, time)
//    records into its own slot range and a descriptor (item, record count)
//    into its work-unit slot (no global atomics); a second kernel
//    (k_conv_sync), one wave per item slot, pushes them to the other copies
//    of the vertex with global atomicMin, marks those dirty and queues their
//    submesh on the next list; a third (k_conv_plan) cuts the next list into
//    work units.  The kernels read only their own list, so results are
//    independent of scheduling (atomic min and "set" flags commute).  Each
//    instance's time window counts from its own start iteration, so a
//    converged instance can be recycled and seeded again mid-run.
//    Iterations are captured in a hipGraph; the host reads one counter per
//    graph launch.
//
//  compat (FP64, reference local update): one workgroup per submesh runs the
//    reference's per-window phases a-g (src/eikonal/cuda_fim.cu:1992-2334)
//    with vertex times and states in LDS; a separate kernel min-merges the
//    copies of interface vertices (sync_step2, cuda_fim.cu:1752-1787).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "host.hpp"
#include "local_solve.hpp"

#define HIPCHECK(x)                                                                   \
  do {                                                                                \
    hipError_t err_ = (x);                                                            \
    if (err_ != hipSuccess)                                                           \
      throw fim::Error(FIM_EDEVICE, std::string(#x) + ": " + hipGetErrorString(err_)); \
  } while (0)

namespace fim {
namespace {

constexpr float F_INF = __builtin_huge_valf();
constexpr int kHardLocalCap = 4096;  // sweeps per activation before requeue

__device__ inline uint64_t lanes_below() {
  const uint32_t lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

__device__ inline bool flag_at(const uint32_t *bits, uint32_t l) {
  return (bits[l >> 5] >> (l & 31)) & 1u;
}

// ---------------------------------------------------------------------------
// converge mode
//
// Work lists are bucketed by submesh: bcnt[s] instances of submesh s are
// queued in bins[s * ninst ...].  A (instance, submesh) pair is queued at most
// once per list (the act flag is set on queueing and cleared when the pair is
// processed), so a bucket never holds more than ninst entries.  After the
// list is complete the plan kernel cuts every bucket into work units of up to
// `unit` instances (plan[u] = s * maxch + chunk, submesh-major) and writes the
// unit count.  Units are dealt out XCD-aware: workgroup b runs on XCD b % 8,
// so the unit range is split into 8 contiguous eighths and workgroups of XCD
// x take units of eighth x first (one static unit each, then from the
// eighth's counter), then help the other eighths.  Instances of one submesh
// thus run on one XCD, back to back, and its element data is fetched into
// that XCD's L2 once.
// ---------------------------------------------------------------------------
struct WorkList {
  uint32_t *bcnt;   // nsub: queued instances per submesh
  uint32_t *bins;   // nsub x ninst instance ids
  uint32_t *plan;   // work units
  uint32_t *cnt;    // [0] units, [1] queued pairs, [2..9] per-XCD unit counters
  uint2 *desc;      // push descriptors, slot u * unit + q for item q of unit u:
                    // (instance * nsub + submesh, push records; 0 = none)
};

constexpr uint32_t kXcd = 8;

__device__ inline void enqueue(const WorkList &l, uint32_t *act, uint32_t inst, uint32_t s,
                               uint32_t nsub, uint32_t ninst) {
  if (atomicExch(&act[(size_t)inst * nsub + s], 1u) == 0u) {
    const uint32_t j = atomicAdd(&l.bcnt[s], 1u);
    l.bins[(size_t)s * ninst + j] = inst;
  }
}

// A work item queues itself for the next iteration.  Its act flag was
// cleared when the item started and only the item's own workgroup can set
// it before the iteration's sync kernel runs, so no exchange is needed.
__device__ inline void requeue(const WorkList &l, uint32_t *act, uint32_t inst, uint32_t s,
                               uint32_t nsub, uint32_t ninst) {
  act[(size_t)inst * nsub + s] = 1u;
  const uint32_t j = atomicAdd(&l.bcnt[s], 1u);
  l.bins[(size_t)s * ninst + j] = inst;
}

// Every item of a list leaves its push descriptor in its unit slot
// (instance * nsub + submesh, push records of its activation, 0 if none), so
// the interface push finds an item's records with one load.
__device__ inline void push_descriptor(const WorkList &l, uint32_t slot, size_t entry, uint32_t n) {
  l.desc[slot] = make_uint2((uint32_t)entry, n);
}

struct UnitRange {
  uint32_t b, e, nb;  // units [b, e); nb workgroups of the grid run on this XCD
};

__device__ inline UnitRange xcd_range(uint32_t U, uint32_t y, uint32_t G) {
  UnitRange r;
  r.b = (uint32_t)(((uint64_t)U * y) / kXcd);
  r.e = (uint32_t)(((uint64_t)U * (y + 1)) / kXcd);
  r.nb = G > y ? (G - y + kXcd - 1) / kXcd : 0u;
  return r;
}

// Iterates the (instance, submesh) pairs of a list for one workgroup; every
// thread of the workgroup calls next() together.  `slot` is one LDS word.
struct UnitCursor {
  const uint32_t *bcnt, *bins, *plan;
  uint32_t *grab;
  uint32_t *slot;
  uint32_t U, G, x, y, d, ninst, unit, maxch;
  UnitRange R;
  uint32_t s, base, n_in, i;
  uint32_t uidx = 0;  // unit of the current item
  uint2 *desc;
  bool first;

  __device__ UnitCursor(const WorkList &l, uint32_t *lds_slot, uint32_t ninst_, uint32_t unit_,
                        uint32_t maxch_)
      : bcnt(l.bcnt), bins(l.bins), plan(l.plan), grab(l.cnt + 2), slot(lds_slot), U(l.cnt[0]),
        G(gridDim.x), x(blockIdx.x % kXcd), y(blockIdx.x % kXcd), d(0), ninst(ninst_),
        unit(unit_), maxch(maxch_), s(0), base(0), n_in(0), i(0), desc(l.desc), first(true) {
    R = xcd_range(U, x, G);
  }

  __device__ bool next(uint32_t &inst, uint32_t &sub) {
    if (++i < n_in) {
      inst = bins[base + i];
      sub = s;
      return true;
    }
    for (;;) {
      uint32_t u;
      if (first) {
        u = R.b + blockIdx.x / kXcd;
        first = false;
      } else {
        // dynamic units follow the R.nb static ones; a plain read of the
        // counter skips exhausted eighths without an atomic
        if (threadIdx.x == 0) {
          const uint32_t dyn = R.b + R.nb;
          uint32_t v = 0xffffffffu;
          if (dyn < R.e &&
              dyn + __hip_atomic_load(&grab[y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < R.e)
            v = dyn + atomicAdd(&grab[y], 1u);
          *slot = v;
        }
        __syncthreads();
        u = *slot;
        __syncthreads();
      }
      if (u < R.e) {
        const uint32_t pu = plan[u];
        s = pu / maxch;
        const uint32_t j = pu - s * maxch;
        n_in = min(unit, bcnt[s] - j * unit);
        base = s * ninst + j * unit;
        i = 0;
        uidx = u;
        // slots of a partial unit that no item fills
        if (threadIdx.x == 0)
          for (uint32_t q = n_in; q < unit; ++q) desc[u * unit + q] = make_uint2(0u, 0u);
        inst = bins[base];
        sub = s;
        return true;
      }
      if (++d == kXcd) return false;
      y = (x + d) % kXcd;
      R = xcd_range(U, y, G);
    }
  }
};

// Builds the work units of a complete list (one workgroup of 1024 threads)
// and clears the buckets of the list the iteration has consumed.
struct PlanArgs {
  const uint32_t *bcnt;
  uint32_t *plan;
  uint32_t *cnt;
  uint32_t *zero_bcnt;  // consumed list, may be null
  uint32_t nsub, unit, maxch;
};

__global__ __launch_bounds__(1024) void k_conv_plan(PlanArgs a) {
  __shared__ uint32_t wsum[16], witems[16];
  __shared__ uint32_t base_sh, items_sh;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) {
    base_sh = 0u;
    items_sh = 0u;
  }
  __syncthreads();
  for (uint32_t s0 = 0; s0 < a.nsub; s0 += 1024) {
    const uint32_t s = s0 + tid;
    uint32_t n = 0;
    if (s < a.nsub) {
      n = a.bcnt[s];
      if (a.zero_bcnt) a.zero_bcnt[s] = 0u;
    }
    const uint32_t ch = (n + a.unit - 1) / a.unit;
    uint32_t v = ch, m = n;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t = __shfl_up(v, off);
      if (lane >= (uint32_t)off) v += t;
      m += __shfl_xor(m, off);
    }
    if (lane == 63) wsum[wave] = v;
    if (lane == 0) witems[wave] = m;
    __syncthreads();
    uint32_t pre = 0, tot = 0, itot = 0;
    for (uint32_t w = 0; w < 16; ++w) {
      if (w < wave) pre += wsum[w];
      tot += wsum[w];
      itot += witems[w];
    }
    const uint32_t off = base_sh + pre + v - ch;
    for (uint32_t j = 0; j < ch; ++j) a.plan[off + j] = s * a.maxch + j;
    __syncthreads();
    if (tid == 0) {
      base_sh += tot;
      items_sh += itot;
    }
    __syncthreads();
  }
  if (tid == 0) {
    a.cnt[0] = base_sh;
    a.cnt[1] = items_sh;
  }
  if (tid < kXcd) a.cnt[2 + tid] = 0u;
}

struct ConvArgs {
  const SubInfo *sub;
  const uint2 *e2n;      // 4 x u16 local ids per element
  const float *geo;      // 6 per element
  const float4 *hgt;     // per element: height of each node above its opposite face
  const uint8_t *ereg;   // region per element
  const float *rscale;   // n_inst x 4 geometry multipliers (1/s^2)
  float *phi;            // n_inst x ncopy
  uint8_t *dirty;        // n_inst x ncopy
  uint32_t *act;         // n_inst x nsub
  float *wake;           // n_inst x nsub: lower bound of the dirty vertex times
  const uint16_t *if_l;  // per interface entry: local copy id
  const uint2 *ifm;      // per submesh, wps words: interface-copy bits of 32
                         // local copies and the interface entries before them
  uint2 *prec;           // push records (entry, time bits): instance i,
                         // submesh s use [i * n_if + ioff(s), + ni(s))
  WorkList cur;          // work list of this iteration
  WorkList nxt;          // work list of the next iteration (requeues)
  unsigned long long *stats;  // activations, sweeps, element evaluations,
                              // iterations, algorithmic bytes
  const uint32_t *iter;  // global iteration counter (advanced by the sync kernel)
  const uint32_t *start; // n_inst: iteration an instance was (re)seeded at; its
                         // window is window * (iter - start + 1)
  const uint32_t *n2e_dsp;  // per copy: incident local elements (staged kernel)
  const uint16_t *n2e_le;
  const uint4 *erec;        // band kernel: 48 B per element (ElemRecord layout)
  const uint32_t *adj_ptr;  // band kernel: per copy, chunk offset << 6 | chunk count
  const uint4 *adj;         // 8 incident local element ids (u16) per chunk, 0xffff padded
  uint32_t nsub, ncopy, ninst;
  float eps;
  const float *winst;    // n_inst: time window added per global iteration of
                         // each instance (us), 0 = off (plain FIM order)
  int32_t max_local;
  uint32_t unit, maxch;  // instances per work unit, units per submesh bucket
  uint32_t n_if, wps;    // interface entries; ifm words per submesh
  uint32_t elcap;        // band kernel: element-list entries per pass
  // LDS carve (bytes)
  uint32_t o_phin, o_flg, o_tch, o_dfr, o_e2n, o_geo, o_list, o_vlist, o_misc, o_wrt, o_ifc;
};

// Writes the interface copies lowered by this activation (bits in tch) as
// push records into the item's slot range and, when there are any, the
// item's push descriptor into its list; the entries are found through the
// interface list (if_l).  Every thread of the workgroup calls it; cnt is one
// LDS word.
template <int BS>
__device__ inline void push_lowered_scan(const ConvArgs &a, const SubInfo &si, uint32_t inst,
                                         size_t entry, uint32_t dslot, const uint32_t *tch,
                                         const float *phi_l, uint32_t *cnt) {
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  if (tid == 0) *cnt = 0u;
  __syncthreads();
  uint2 *rec = a.prec + (size_t)inst * a.n_if + si.ioff;
  for (uint32_t j0 = 0; j0 < si.ni; j0 += BS) {
    const uint32_t j = j0 + tid;
    uint32_t l = 0;
    bool p = false;
    if (j < si.ni) {
      l = a.if_l[si.ioff + j];
      p = flag_at(tch, l);
    }
    const uint64_t m = __ballot(p);
    if (m == 0) continue;  // uniform across the wave
    uint32_t wb = 0;
    if (lane == 0) wb = atomicAdd(cnt, (uint32_t)__popcll(m));
    wb = __shfl(wb, 0);
    if (p) rec[wb + __popcll(m & lanes_below())] = make_uint2(si.ioff + j, __float_as_uint(phi_l[l]));
  }
  __syncthreads();
  if (tid == 0) push_descriptor(a.cur, dslot, entry, *cnt);
}

// One workgroup per work item (instance, submesh).  STAGED copies the
// submesh's elements (nodes + geometry) into LDS and finds the active
// elements by scanning them every sweep; otherwise element data stays in
// global memory (L2 / MALL resident) and the active elements are marked
// from the flagged vertices through the vertex -> element adjacency, so an
// activation costs in proportion to the band of the submesh it touches.
template <int BS, bool STAGED>
__global__ __launch_bounds__(BS) void k_conv_compute(ConvArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  float *phi_l = reinterpret_cast<float *>(lds);
  uint32_t *phi_n = reinterpret_cast<uint32_t *>(lds + a.o_phin);
  uint32_t *flg = reinterpret_cast<uint32_t *>(lds + a.o_flg);
  uint32_t *tch = reinterpret_cast<uint32_t *>(lds + a.o_tch);
  uint32_t *dfr = reinterpret_cast<uint32_t *>(lds + a.o_dfr);
  uint2 *el = reinterpret_cast<uint2 *>(lds + a.o_e2n);      // STAGED
  float *gl = reinterpret_cast<float *>(lds + a.o_geo);       // STAGED
  uint32_t *emark = reinterpret_cast<uint32_t *>(lds + a.o_e2n);  // !STAGED
  uint32_t *eseen = reinterpret_cast<uint32_t *>(lds + a.o_geo);  // !STAGED
  uint16_t *elist = reinterpret_cast<uint16_t *>(lds + a.o_list);
  uint16_t *vlist = reinterpret_cast<uint16_t *>(lds + a.o_vlist);  // !STAGED
  uint32_t *misc = reinterpret_cast<uint32_t *>(lds + a.o_misc);
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  UnitCursor cursor(a.cur, &misc[4], a.ninst, a.unit, a.maxch);
  if (blockIdx.x == 0 && tid == 0 && cursor.U > 0) atomicAdd(&a.stats[3], 1ull);
  unsigned long long sweeps = 0, evals = 0, acts = 0, abytes = 0, vupd = 0;
  // time window of this iteration: flagged vertices later than W wait
  const uint32_t iter_now = *a.iter;

  uint32_t inst, s;
  while (cursor.next(inst, s)) {
    const size_t entry = (size_t)inst * a.nsub + s;
    const float wi = a.winst[inst];
    const float W = wi > 0.0f ? wi * (float)(iter_now - a.start[inst] + 1u) : F_INF;
    const SubInfo si = a.sub[s];
    const size_t cb = (size_t)inst * a.ncopy + si.coff;
    const uint32_t nn = si.nn, ne = si.ne;
    const uint32_t nv64 = (nn + 63u) & ~63u;
    const uint32_t new32 = (ne + 31u) >> 5;
    if (tid == 0) a.act[entry] = 0u;
    // nothing in this submesh can act before the window reaches its wake
    // time: keep it queued without staging anything
    if (a.wake[entry] > W) {
      if (tid == 0) {
        requeue(a.nxt, a.act, inst, s, a.nsub, a.ninst);
        push_descriptor(a.cur, cursor.uidx * a.unit + cursor.i, entry, 0u);
      }
      continue;
    }
    // stage vertex times and dirty flags
    for (uint32_t l0 = 0; l0 < nv64; l0 += BS) {
      const uint32_t l = l0 + tid;
      bool d = false, late = false;
      if (l < nn) {
        const float v = a.phi[cb + l];
        phi_l[l] = v;
        phi_n[l] = __float_as_uint(v);
        d = a.dirty[cb + l] != 0;
        if (d) a.dirty[cb + l] = 0;
        late = d && v > W;
        d = d && !late;
      }
      const uint64_t m = __ballot(d), md = __ballot(late);
      const uint32_t wb = l0 + wave * 64;
      if (wb < nv64) {
        if (lane == 0) flg[wb >> 5] = (uint32_t)m;
        if (lane == 1) flg[(wb >> 5) + 1] = (uint32_t)(m >> 32);
        if (lane == 2) tch[wb >> 5] = 0u;
        if (lane == 3) tch[(wb >> 5) + 1] = 0u;
        if (lane == 4) dfr[wb >> 5] = (uint32_t)md;
        if (lane == 5) dfr[(wb >> 5) + 1] = (uint32_t)(md >> 32);
      }
    }
    const float *rs = a.rscale + 4 * (size_t)inst;
    if (STAGED) {
      // stage elements (geometry scaled by the instance's regional factor)
      for (uint32_t e = tid; e < ne; e += BS) {
        const size_t ge = (size_t)si.eoff + e;
        el[e] = a.e2n[ge];
        const float sc = rs[a.ereg[ge] & 3];
        const float2 *g2 = reinterpret_cast<const float2 *>(a.geo + 6 * ge);
        const float2 x0 = g2[0], x1 = g2[1], x2 = g2[2];
        float *o = gl + 6 * e;
        o[0] = x0.x * sc;
        o[1] = x0.y * sc;
        o[2] = x1.x * sc;
        o[3] = x1.y * sc;
        o[4] = x2.x * sc;
        o[5] = x2.y * sc;
      }
    } else {
      for (uint32_t k = tid; k < new32; k += BS) {
        emark[k] = 0u;
        eseen[k] = 0u;
      }
    }
    __syncthreads();

    int it = 0;
    bool capped = false;
    for (;;) {
      if (tid == 0) {
        misc[0] = 0u;
        misc[2] = 0u;
      }
      __syncthreads();
      if (STAGED) {
        // active elements: any node flagged (wave ballot + per-wave base)
        for (uint32_t e0 = 0; e0 < ne; e0 += BS) {
          const uint32_t e = e0 + tid;
          bool on = false;
          if (e < ne) {
            const uint2 q = el[e];
            on = flag_at(flg, q.x & 0xffffu) || flag_at(flg, q.x >> 16) ||
                 flag_at(flg, q.y & 0xffffu) || flag_at(flg, q.y >> 16);
          }
          const uint64_t m = __ballot(on);
          uint32_t base = 0;
          if (lane == 0 && m) base = atomicAdd(&misc[0], (uint32_t)__popcll(m));
          base = __shfl(base, 0);
          if (on) elist[base + __popcll(m & lanes_below())] = (uint16_t)e;
        }
        __syncthreads();
      } else {
        // flagged vertices -> compact list
        for (uint32_t l0 = 0; l0 < nv64; l0 += BS) {
          const uint32_t l = l0 + tid;
          const bool on = l < nn && flag_at(flg, l);
          const uint64_t m = __ballot(on);
          uint32_t base = 0;
          if (lane == 0 && m) base = atomicAdd(&misc[2], (uint32_t)__popcll(m));
          base = __shfl(base, 0);
          if (on) vlist[base + __popcll(m & lanes_below())] = (uint16_t)l;
        }
        __syncthreads();
        const uint32_t nsrc = misc[2];
        // mark the elements incident to them
        for (uint32_t i = tid; i < nsrc; i += BS) {
          const uint32_t c = si.coff + vlist[i];
          const uint32_t kb = a.n2e_dsp[c], ke = a.n2e_dsp[c + 1];
          for (uint32_t k = kb; k < ke; ++k) {
            const uint32_t le = a.n2e_le[k];
            atomicOr(&emark[le >> 5], 1u << (le & 31));
          }
        }
        __syncthreads();
        // marked elements -> compact list (and clear the marks)
        for (uint32_t k0 = 0; k0 < new32; k0 += BS) {
          const uint32_t k = k0 + tid;
          uint32_t bits = 0;
          if (k < new32) {
            bits = emark[k];
            emark[k] = 0u;
            eseen[k] |= bits;
          }
          const uint32_t cnt = __popc(bits);
          uint32_t base = cnt ? atomicAdd(&misc[0], cnt) : 0u;
          while (bits) {
            const uint32_t b = __ffs(bits) - 1;
            bits &= bits - 1;
            elist[base++] = (uint16_t)(k * 32 + b);
          }
        }
        __syncthreads();
      }
      const uint32_t nact = misc[0];
      if (nact == 0) break;
      if ((a.max_local > 0 && it >= a.max_local) || it >= kHardLocalCap) {
        capped = true;
        break;
      }
      ++it;
      evals += nact;
      // candidates of the active elements
      for (uint32_t i = tid; i < nact; i += BS) {
        const uint32_t e = elist[i];
        uint2 q;
        float g[6];
        if (STAGED) {
          q = el[e];
#pragma unroll
          for (int k = 0; k < 6; ++k) g[k] = gl[6 * e + k];
        } else {
          const size_t ge = (size_t)si.eoff + e;
          q = a.e2n[ge];
          const float sc = rs[a.ereg[ge] & 3];
          const float2 *g2 = reinterpret_cast<const float2 *>(a.geo + 6 * ge);
          const float2 x0 = g2[0], x1 = g2[1], x2 = g2[2];
          g[0] = x0.x * sc;
          g[1] = x0.y * sc;
          g[2] = x1.x * sc;
          g[3] = x1.y * sc;
          g[4] = x2.x * sc;
          g[5] = x2.y * sc;
        }
        const uint32_t nd[4] = {q.x & 0xffffu, q.x >> 16, q.y & 0xffffu, q.y >> 16};
        float ph[4];
        int f[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          ph[k] = phi_l[nd[k]];
          f[k] = flag_at(flg, nd[k]) ? 1 : 0;
        }
        const int nf = f[0] + f[1] + f[2] + f[3];
        // every candidate of a target is at least the earliest other node
        // plus the target's metric height above the opposite face: evaluate
        // only targets with a flagged other node that can still improve
        const size_t gh = (size_t)si.eoff + e;
        const float hs = csqrt(rs[a.ereg[gh] & 3]);
        const float4 hh = a.hgt[gh];
        const float h[4] = {hh.x * hs, hh.y * hs, hh.z * hs, hh.w * hs};
        if (nf - f[0] > 0 && ph[0] > fminf(ph[1], fminf(ph[2], ph[3])) + h[0]) {
          const float c = cand_exact<0, float>(g, ph, F_INF);
          ++vupd;
          if (c < ph[0]) atomicMin(&phi_n[nd[0]], __float_as_uint(c));
        }
        if (nf - f[1] > 0 && ph[1] > fminf(ph[0], fminf(ph[2], ph[3])) + h[1]) {
          const float c = cand_exact<1, float>(g, ph, F_INF);
          ++vupd;
          if (c < ph[1]) atomicMin(&phi_n[nd[1]], __float_as_uint(c));
        }
        if (nf - f[2] > 0 && ph[2] > fminf(ph[0], fminf(ph[1], ph[3])) + h[2]) {
          const float c = cand_exact<2, float>(g, ph, F_INF);
          ++vupd;
          if (c < ph[2]) atomicMin(&phi_n[nd[2]], __float_as_uint(c));
        }
        if (nf - f[3] > 0 && ph[3] > fminf(ph[0], fminf(ph[1], ph[2])) + h[3]) {
          const float c = cand_exact<3, float>(g, ph, F_INF);
          ++vupd;
          if (c < ph[3]) atomicMin(&phi_n[nd[3]], __float_as_uint(c));
        }
      }
      __syncthreads();
      // accept: new flags = significant decreases (reference criterion
      // |old-new| >= eps (1 + (old+new)/2), cuda_fim.cu:2154)
      for (uint32_t l0 = 0; l0 < nv64; l0 += BS) {
        const uint32_t l = l0 + tid;
        bool sig = false, dec = false, late = false;
        if (l < nn) {
          const float ov = phi_l[l];
          const float nv = __uint_as_float(phi_n[l]);
          if (nv < ov) {
            phi_l[l] = nv;
            dec = true;
            sig = (ov - nv) >= a.eps * (1.0f + 0.5f * (ov + nv));
            late = sig && nv > W;
            sig = sig && !late;
          }
        }
        const uint64_t m = __ballot(sig), mt = __ballot(dec), md = __ballot(late);
        const uint32_t wb = l0 + wave * 64;
        if (wb < nv64) {
          if (lane == 0) flg[wb >> 5] = (uint32_t)m;
          if (lane == 1) flg[(wb >> 5) + 1] = (uint32_t)(m >> 32);
          if (lane == 2) tch[wb >> 5] |= (uint32_t)mt;
          if (lane == 3) tch[(wb >> 5) + 1] |= (uint32_t)(mt >> 32);
          // a deferred vertex that is now a source leaves the deferred set
          if (lane == 4) dfr[wb >> 5] = (dfr[wb >> 5] & ~(uint32_t)m) | (uint32_t)md;
          if (lane == 5) dfr[(wb >> 5) + 1] = (dfr[(wb >> 5) + 1] & ~(uint32_t)(m >> 32)) | (uint32_t)(md >> 32);
        }
      }
      __syncthreads();
    }
    sweeps += it;
    acts += 1;
    // algorithmic bytes of the activation: copies read and written (phi 4+4,
    // flag 1+1), interface entries (index 4, read 4, write 4) and the
    // element data used (nodes 8 + geometry 24) -- all elements when staged,
    // the distinct elements evaluated otherwise
    uint32_t used = ne;
    if (!STAGED) {
      if (tid == 0) misc[3] = 0u;
      __syncthreads();
      uint32_t c = 0;
      for (uint32_t k = tid; k < new32; k += BS) c += __popc(eseen[k]);
      if (c) atomicAdd(&misc[3], c);
      __syncthreads();
      used = misc[3];
    }
    abytes += 32ull * used + 10ull * nn + 12ull * si.ni;
    // write back decreased copies and push the lowered interface copies.
    // Deferred vertices (and sources left by the sweep cap) stay dirty and
    // the submesh is queued for the next iteration.
    if (tid == 0) misc[1] = __float_as_uint(F_INF);
    __syncthreads();
    for (uint32_t l = tid; l < nn; l += BS) {
      if (flag_at(tch, l)) a.phi[cb + l] = phi_l[l];
      if (flag_at(dfr, l) || (capped && flag_at(flg, l))) {
        a.dirty[cb + l] = 1;
        atomicMin(&misc[1], __float_as_uint(phi_l[l]));
      }
    }
    push_lowered_scan<BS>(a, si, inst, entry, cursor.uidx * a.unit + cursor.i, tch, phi_l, &misc[6]);
    __syncthreads();
    if (tid == 0) {
      const float wk = __uint_as_float(misc[1]);
      a.wake[entry] = wk;
      if (wk < F_INF) requeue(a.nxt, a.act, inst, s, a.nsub, a.ninst);
    }
    __syncthreads();
  }
  if (tid == 0 && acts) {
    atomicAdd(&a.stats[0], acts);
    atomicAdd(&a.stats[1], sweeps);
    atomicAdd(&a.stats[2], evals);
    atomicAdd(&a.stats[4], abytes);
  }
  for (int off = 32; off > 0; off >>= 1) vupd += __shfl_xor(vupd, off);
  if (lane == 0 && vupd) atomicAdd(&a.stats[5], vupd);
}

// Element record of the band kernel, three 16-byte words per element:
//   [0] nodes 0,1 (u16 each; bit 15 and bit 31 carry the region id's bits 0
//       and 1), nodes 2,3 (u16 each), geometry e12 e13 (f32)
//   [1] geometry e23 e14 e24 e34 (f32)
//   [2] node heights h0..h3 (f32)
// Local copy ids are below 2^15, so the two region bits are free.
constexpr uint32_t kAdjPad = 0xffffffffu;
// (element, target) pair ring of one wave in the band kernel: up to 63 pairs
// left from earlier rounds plus 4 targets of 64 elements
constexpr uint32_t kPairRing = 320;

// marks the (up to two) local elements of one adjacency word in the LDS
// bitmask; the atomics return nothing, the list is built from the bitmask
__device__ inline void mark_word(uint32_t *emark, uint32_t w) {
  const uint32_t lo = w & 0xffffu, hi = w >> 16;
  if (lo != 0xffffu) atomicOr(&emark[lo >> 5], 1u << (lo & 31));
  if (hi != 0xffffu) atomicOr(&emark[hi >> 5], 1u << (hi & 31));
}

__device__ inline void mark_chunk(uint32_t *emark, const uint4 q) {
  mark_word(emark, q.x);
  mark_word(emark, q.y);
  mark_word(emark, q.z);
  mark_word(emark, q.w);
}

// marks the incident elements of copy c (global copy id): up to four 8-entry
// chunks of the padded adjacency are loaded before any is used
__device__ inline void mark_incident(const ConvArgs &a, uint32_t c, uint32_t *emark) {
  const uint32_t ap = a.adj_ptr[c];
  const uint4 *ch = a.adj + (ap >> 6);
  const uint32_t nch = ap & 63u;
  const uint4 pad = make_uint4(kAdjPad, kAdjPad, kAdjPad, kAdjPad);
  const uint4 q0 = ch[0];
  const uint4 q1 = nch > 1 ? ch[1] : pad;
  const uint4 q2 = nch > 2 ? ch[2] : pad;
  const uint4 q3 = nch > 3 ? ch[3] : pad;
  mark_chunk(emark, q0);
  mark_chunk(emark, q1);
  mark_chunk(emark, q2);
  mark_chunk(emark, q3);
  for (uint32_t k = 4; k < nch; ++k) mark_chunk(emark, ch[k]);
}

// sum over the wave of n < 2^BITS, one ballot per bit; every lane of the
// wave must be active
template <int BITS>
__device__ inline uint32_t wave_sum_small(uint32_t n) {
  uint32_t s = 0;
#pragma unroll
  for (int b = 0; b < BITS; ++b) s += (uint32_t)__popcll(__ballot((n >> b) & 1u)) << b;
  return s;
}

// appends the lanes with `pred` (vertex l) to list[*cnt ...], one LDS atomic
// per wave
__device__ inline void wave_append(bool pred, uint32_t l, uint16_t *list, uint32_t *cnt) {
  const uint64_t m = __ballot(pred);
  uint32_t base = 0;
  if ((threadIdx.x & 63) == 0 && m) base = atomicAdd(cnt, (uint32_t)__popcll(m));
  base = __builtin_amdgcn_readfirstlane(base);  // every lane of the wave is active
  if (pred) list[base + __popcll(m & lanes_below())] = (uint16_t)l;
}

// Unstaged converge kernel.  Per sweep it touches only the band that is
// active, in three steps separated by barriers:
//   mark: the sources (a compact list) mark their incident elements through
//     the vertex -> element adjacency in an LDS bitmask (atomics without
//     return), then one thread per bitmask word lists the marked elements in
//     id order and clears the word;
//   evaluate: the listed elements compute candidates from L2-resident
//     element records; LDS atomicMin winners set a "written" bit;
//   accept: each written vertex takes its new time; a significant decrease
//     (reference criterion, cuda_fim.cu:2154) makes it a source of the next
//     sweep, or a deferred vertex if it lies beyond the window.
// Element data never enters LDS.  PAIRS evaluates (element, target) pairs
// in full rounds of 64 instead of an element's four targets per lane (local
// element ids below 2^14).  WPE is the number of waves per SIMD the
// register budget is sized for (512 / WPE VGPRs): the launch picks the one
// that the LDS carve allows anyway, so registers left over by a large
// submesh's LDS footprint are used instead of spilled.
template <int BS, int WPE, bool PAIRS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_conv_band(ConvArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  float *phi_l = reinterpret_cast<float *>(lds);
  uint32_t *phi_n = reinterpret_cast<uint32_t *>(lds + a.o_phin);
  uint32_t *flg = reinterpret_cast<uint32_t *>(lds + a.o_flg);  // sources of the sweep
  uint32_t *tch = reinterpret_cast<uint32_t *>(lds + a.o_tch);  // decreased this activation
  uint32_t *dfr = reinterpret_cast<uint32_t *>(lds + a.o_dfr);  // flagged, beyond the window
  uint32_t *wrt = reinterpret_cast<uint32_t *>(lds + a.o_wrt);  // lowered in this sweep
  uint32_t *ifc = reinterpret_cast<uint32_t *>(lds + a.o_ifc);  // interface copies
  uint32_t *ipre = ifc + a.wps;  // interface entries before each word
  uint32_t *emark = reinterpret_cast<uint32_t *>(lds + a.o_e2n);
  uint32_t *eseen = reinterpret_cast<uint32_t *>(lds + a.o_geo);
  uint16_t *elist = reinterpret_cast<uint16_t *>(lds + a.o_list);
  uint16_t *vlist = reinterpret_cast<uint16_t *>(lds + a.o_vlist);
  // misc: [0] listed elements, [2] listed sources, [3] distinct elements,
  // [4] work cursor, [5] earliest deferred time, [6] push records,
  // [8..11] the instance's region factors
  uint32_t *misc = reinterpret_cast<uint32_t *>(lds + a.o_misc);
  const float *rsl = reinterpret_cast<const float *>(misc + 8);
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  UnitCursor cursor(a.cur, &misc[4], a.ninst, a.unit, a.maxch);
  if (blockIdx.x == 0 && tid == 0 && cursor.U > 0) atomicAdd(&a.stats[3], 1ull);
  unsigned long long sweeps = 0, evals = 0, acts = 0, abytes = 0, vupd = 0;
  const uint32_t iter_now = *a.iter;

  uint32_t inst, s;
  while (cursor.next(inst, s)) {
    const size_t entry = (size_t)inst * a.nsub + s;
    const float wi = a.winst[inst];
    const float W = wi > 0.0f ? wi * (float)(iter_now - a.start[inst] + 1u) : F_INF;
    const SubInfo si = a.sub[s];
    const size_t cb = (size_t)inst * a.ncopy + si.coff;
    const uint32_t nn = si.nn, ne = si.ne;
    const uint32_t nv64 = (nn + 63u) & ~63u;
    const uint32_t new32 = (ne + 31u) >> 5;  // element bitmask words
    if (tid == 0) a.act[entry] = 0u;
    if (a.wake[entry] > W) {
      if (tid == 0) {
        requeue(a.nxt, a.act, inst, s, a.nsub, a.ninst);
        push_descriptor(a.cur, cursor.uidx * a.unit + cursor.i, entry, 0u);
      }
      continue;
    }
    for (uint32_t k = tid; k < new32; k += BS) {
      emark[k] = 0u;
      eseen[k] = 0u;
    }
    if (tid == 0) {
      misc[0] = 0u;
      misc[2] = 0u;
      misc[3] = 0u;
      misc[5] = __float_as_uint(F_INF);
      misc[6] = 0u;
    }
    if (tid < 4) misc[8 + tid] = __float_as_uint(a.rscale[4 * (size_t)inst + tid]);
    __syncthreads();
    // stage vertex times; dirty flags split into sources and deferred
    for (uint32_t l0 = 0; l0 < nv64; l0 += BS) {
      const uint32_t l = l0 + tid;
      const uint32_t wb = l0 + wave * 64;
      if (wb < nv64) {
        bool d = false, late = false;
        if (l < nn) {
          const float v = a.phi[cb + l];
          phi_l[l] = v;
          phi_n[l] = __float_as_uint(v);
          d = a.dirty[cb + l] != 0;
          if (d) a.dirty[cb + l] = 0;
          late = d && v > W;
          d = d && !late;
        }
        const uint64_t m = __ballot(d), md = __ballot(late);
        const uint32_t k = wb >> 5;
        if (lane == 0) {
          flg[k] = (uint32_t)m;
          flg[k + 1] = (uint32_t)(m >> 32);
          dfr[k] = (uint32_t)md;
          dfr[k + 1] = (uint32_t)(md >> 32);
        }
        if (lane == 1) {
          tch[k] = 0u;
          tch[k + 1] = 0u;
          wrt[k] = 0u;
          wrt[k + 1] = 0u;
        }
        if (lane == 2 || lane == 3) {
          const uint2 w = a.ifm[(size_t)s * a.wps + k + (lane - 2)];
          ifc[k + lane - 2] = w.x;
          ipre[k + lane - 2] = w.y;
        }
        wave_append(d, l, vlist, &misc[2]);
      }
    }
    __syncthreads();

    uint32_t it = 0;
    bool capped = false;
    for (;;) {
      const uint32_t nsrc = misc[2];
      if (nsrc == 0) break;
      if ((a.max_local > 0 && it >= (uint32_t)a.max_local) || it >= (uint32_t)kHardLocalCap) {
        capped = true;
        break;
      }
      ++it;
      // mark the elements incident to the sources, then list the marked
      // elements in id order (clearing the marks, one word per thread)
      for (uint32_t i = tid; i < nsrc; i += BS) mark_incident(a, si.coff + vlist[i], emark);
      __syncthreads();
      // element-list passes of at most elcap elements: a word's elements
      // beyond the cap stay marked for the next pass (the sweep still
      // evaluates every marked element against the same times)
      for (;;) {
        for (uint32_t k = tid; k < new32; k += BS) {
          // list offsets: a wave prefix sum of the word counts (<= 32, six
          // bits) from one ballot per bit, one LDS reservation per wave
          uint32_t rest = emark[k];
          const uint32_t n = __popc(rest);
          uint32_t pre = 0, tot = 0;
#pragma unroll
          for (int b = 0; b < 6; ++b) {
            const uint64_t m = __ballot((n >> b) & 1u);
            pre += (uint32_t)__popcll(m & lanes_below()) << b;
            tot += (uint32_t)__popcll(m) << b;
          }
          uint32_t base = 0;
          if (lane == 0 && tot) base = atomicAdd(&misc[0], tot);  // lane 0 is active here
          uint32_t pos = __builtin_amdgcn_readfirstlane(base) + pre;
          if (rest && pos < a.elcap) {
            uint32_t take = 0;
            while (rest && pos < a.elcap) {
              const uint32_t b = __ffs(rest) - 1;
              rest &= rest - 1;
              take |= 1u << b;
              elist[pos++] = (uint16_t)(k * 32 + b);
            }
            emark[k] = rest;
            eseen[k] |= take;
          }
        }
        __syncthreads();
        const uint32_t listed = misc[0];
        const uint32_t nact = min(listed, a.elcap);
        evals += nact;
        if (tid == 0) misc[2] = 0u;  // every thread has read the source count
        // candidates of the listed elements
        if constexpr (PAIRS) {
          // (element, target) pairs: a pass over 64 listed elements per wave
          // keeps the targets that pass the bound test and queues them in a
          // wave-private ring (LDS, in the source list's space, idle until the
          // accept step); full rounds of 64 pairs then evaluate one target
          // per lane, so no lane idles on a pruned target
          uint16_t *pq = vlist + wave * kPairRing;
          uint32_t qh = 0, qn = 0;  // ring head and fill, uniform across the wave
          for (uint32_t c0 = wave * 64;; c0 += BS) {
            const bool more = c0 < nact;  // uniform across the wave
            if (more) {
              const uint32_t i = c0 + lane;
              uint32_t tm = 0, e = 0;
              if (i < nact) {
                e = elist[i];
                const uint4 *rec = a.erec + 3 * ((size_t)si.eoff + e);
                const uint4 r0 = rec[0], r2 = rec[2];
                const float hs = csqrt(rsl[((r0.x >> 15) & 1u) | ((r0.x >> 30) & 2u)]);
                const float h[4] = {__uint_as_float(r2.x) * hs, __uint_as_float(r2.y) * hs,
                                    __uint_as_float(r2.z) * hs, __uint_as_float(r2.w) * hs};
                const uint32_t nd[4] = {r0.x & 0x7fffu, (r0.x >> 16) & 0x7fffu, r0.y & 0xffffu,
                                        r0.y >> 16};
                float ph[4];
                int f[4];
  #pragma unroll
                for (int k = 0; k < 4; ++k) {
                  ph[k] = phi_l[nd[k]];
                  f[k] = flag_at(flg, nd[k]) ? 1 : 0;
                }
                const int nf = f[0] + f[1] + f[2] + f[3];
                // targets with a source among the other nodes whose bound
                // (earliest other node + metric height) is still below their time
                tm |= (nf - f[0] > 0 && ph[0] > fminf(ph[1], fminf(ph[2], ph[3])) + h[0]) ? 1u : 0u;
                tm |= (nf - f[1] > 0 && ph[1] > fminf(ph[0], fminf(ph[2], ph[3])) + h[1]) ? 2u : 0u;
                tm |= (nf - f[2] > 0 && ph[2] > fminf(ph[0], fminf(ph[1], ph[3])) + h[2]) ? 4u : 0u;
                tm |= (nf - f[3] > 0 && ph[3] > fminf(ph[0], fminf(ph[1], ph[2])) + h[3]) ? 8u : 0u;
              }
  #pragma unroll
              for (uint32_t T = 0; T < 4; ++T) {
                const bool p = (tm >> T) & 1u;
                const uint64_t m = __ballot(p);
                if (p) {
                  uint32_t pos = qh + qn + (uint32_t)__popcll(m & lanes_below());
                  if (pos >= kPairRing) pos -= kPairRing;
                  pq[pos] = (uint16_t)(e | (T << 14));
                }
                qn += (uint32_t)__popcll(m);
              }
              __builtin_amdgcn_wave_barrier();
            }
            while (qn >= 64u || (!more && qn > 0u)) {
              const uint32_t take = qn < 64u ? qn : 64u;
              if (lane < take) {
                uint32_t pos = qh + lane;
                if (pos >= kPairRing) pos -= kPairRing;
                const uint32_t code = pq[pos];
                const uint32_t e = code & 0x3fffu, T = code >> 14;
                const uint4 *rec = a.erec + 3 * ((size_t)si.eoff + e);
                const uint4 r0 = rec[0], r1 = rec[1];
                const float sc = rsl[((r0.x >> 15) & 1u) | ((r0.x >> 30) & 2u)];
                const float g[6] = {__uint_as_float(r0.z) * sc, __uint_as_float(r0.w) * sc,
                                    __uint_as_float(r1.x) * sc, __uint_as_float(r1.y) * sc,
                                    __uint_as_float(r1.z) * sc, __uint_as_float(r1.w) * sc};
                const uint32_t nd[4] = {r0.x & 0x7fffu, (r0.x >> 16) & 0x7fffu, r0.y & 0xffffu,
                                        r0.y >> 16};
                // relabel so that the target is node 3 (others T+1, T+2, T+3):
                // the same candidate as for node position T
                const bool t1 = T & 1u, t2 = T & 2u;
                auto sel = [&](float x0, float x1, float x2, float x3) {
                  return t2 ? (t1 ? x3 : x2) : (t1 ? x1 : x0);
                };
                auto seli = [&](uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3) {
                  return t2 ? (t1 ? x3 : x2) : (t1 ? x1 : x0);
                };
                // edge (i,j) of the relabelled element for T = 0, 1, 2, 3
                const float gp[6] = {sel(g[2], g[5], g[3], g[0]), sel(g[4], g[1], g[4], g[1]),
                                     sel(g[5], g[3], g[0], g[2]), sel(g[0], g[2], g[5], g[3]),
                                     sel(g[1], g[4], g[1], g[4]), sel(g[3], g[0], g[2], g[5])};
                const uint32_t n1 = seli(nd[1], nd[2], nd[3], nd[0]);
                const uint32_t n2 = seli(nd[2], nd[3], nd[0], nd[1]);
                const uint32_t n3 = seli(nd[3], nd[0], nd[1], nd[2]);
                const uint32_t nt = seli(nd[0], nd[1], nd[2], nd[3]);
                const float pp[4] = {phi_l[n1], phi_l[n2], phi_l[n3], phi_l[nt]};
                float sq[6], rc[6];
  #pragma unroll
                for (int k = 0; k < 3; ++k) {
                  rc[k] = __builtin_amdgcn_rcpf(gp[k]);
                  sq[k + 3] = csqrt(gp[k + 3]);
                  sq[k] = 0.0f;
                  rc[k + 3] = 0.0f;
                }
                const float c = cand_exact_sh<3, float>(gp, sq, rc, pp, F_INF);
                ++vupd;
                if (c < pp[3]) {
                  const uint32_t cbits = __float_as_uint(c);
                  // below the staged time: the minimum ends below it too, so the
                  // written bit needs no returned value
                  atomicMin(&phi_n[nt], cbits);
                  atomicOr(&wrt[nt >> 5], 1u << (nt & 31));
                }
              }
              qh += take;
              if (qh >= kPairRing) qh -= kPairRing;
              qn -= take;
              __builtin_amdgcn_wave_barrier();
            }
            if (!more) break;
          }
        } else {
          for (uint32_t i = tid; i < nact; i += BS) {
            const uint32_t e = elist[i];
            const uint4 *rec = a.erec + 3 * ((size_t)si.eoff + e);
            const uint4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
            const float sc = rsl[((r0.x >> 15) & 1u) | ((r0.x >> 30) & 2u)];
            const float g[6] = {__uint_as_float(r0.z) * sc, __uint_as_float(r0.w) * sc,
                                __uint_as_float(r1.x) * sc, __uint_as_float(r1.y) * sc,
                                __uint_as_float(r1.z) * sc, __uint_as_float(r1.w) * sc};
            // square roots (edge candidates) and reciprocals (triangle bases) of
            // the six squared lengths, shared by the element's four targets
            float sq[6], rc[6];
    #pragma unroll
            for (int k = 0; k < 6; ++k) {
              sq[k] = csqrt(g[k]);
              rc[k] = __builtin_amdgcn_rcpf(g[k]);
            }
            const float hs = csqrt(sc);
            const float h[4] = {__uint_as_float(r2.x) * hs, __uint_as_float(r2.y) * hs,
                                __uint_as_float(r2.z) * hs, __uint_as_float(r2.w) * hs};
            const uint32_t nd[4] = {r0.x & 0x7fffu, (r0.x >> 16) & 0x7fffu, r0.y & 0xffffu,
                                    r0.y >> 16};
            float ph[4];
            int f[4];
    #pragma unroll
            for (int k = 0; k < 4; ++k) {
              ph[k] = phi_l[nd[k]];
              f[k] = flag_at(flg, nd[k]) ? 1 : 0;
            }
            const int nf = f[0] + f[1] + f[2] + f[3];
            // targets with a source among the other nodes whose bound (earliest
            // other node + metric height) is still below their time
    #define FIM_TARGET(T, A, B, C)                                                        \
            if (nf - f[T] > 0 && ph[T] > fminf(ph[A], fminf(ph[B], ph[C])) + h[T]) {         \
              const float c = cand_exact_sh<T, float>(g, sq, rc, ph, F_INF);                 \
              ++vupd;                                                                         \
              if (c < ph[T]) {                                                                \
                const uint32_t cbits = __float_as_uint(c);                                    \
                if (cbits < atomicMin(&phi_n[nd[T]], cbits))                                  \
                  atomicOr(&wrt[nd[T] >> 5], 1u << (nd[T] & 31));                             \
              }                                                                               \
            }
            FIM_TARGET(0, 1, 2, 3)
            FIM_TARGET(1, 0, 2, 3)
            FIM_TARGET(2, 0, 1, 3)
            FIM_TARGET(3, 0, 1, 2)
    #undef FIM_TARGET
          }
        }
        __syncthreads();
        if (tid == 0) misc[0] = 0u;  // every thread has read the element count
        if (listed <= a.elcap) break;  // uniform: read after a barrier
        __syncthreads();
      }
      // accept the written vertices: the sources of the next sweep
      for (uint32_t l0 = 0; l0 < nv64; l0 += BS) {
        const uint32_t l = l0 + tid;
        const uint32_t wb = l0 + wave * 64;
        if (wb < nv64) {
          const uint32_t k = wb >> 5;
          // a block of 64 copies with nothing written has no sources
          if ((wrt[k] | wrt[k + 1]) == 0u) {  // uniform: same LDS words
            if (lane == 0) {
              flg[k] = 0u;
              flg[k + 1] = 0u;
            }
            continue;
          }
          bool sg = false, lt = false, dc = false;
          if (l < nn && ((wrt[l >> 5] >> (l & 31)) & 1u)) {
            const float ov = phi_l[l];
            const float nv = __uint_as_float(phi_n[l]);
            if (nv < ov) {
              phi_l[l] = nv;
              dc = true;
              if ((ov - nv) >= a.eps * (1.0f + 0.5f * (ov + nv))) {
                lt = nv > W;
                sg = !lt;
              }
            }
          }
          const uint64_t msg = __ballot(sg), mlt = __ballot(lt), mdc = __ballot(dc);
          if (lane == 0) {
            flg[k] = (uint32_t)msg;
            flg[k + 1] = (uint32_t)(msg >> 32);
            wrt[k] = 0u;
            wrt[k + 1] = 0u;
          }
          if (lane == 1) {
            tch[k] |= (uint32_t)mdc;
            tch[k + 1] |= (uint32_t)(mdc >> 32);
          }
          if (lane == 2) {
            dfr[k] = (dfr[k] & ~(uint32_t)msg) | (uint32_t)mlt;
            dfr[k + 1] = (dfr[k + 1] & ~(uint32_t)(msg >> 32)) | (uint32_t)(mlt >> 32);
          }
          wave_append(sg, l, vlist, &misc[2]);
        }
      }
      __syncthreads();
    }
    sweeps += it;
    acts += 1;
    // algorithmic bytes: copies (phi 4+4, flag 1+1), interface entries
    // (4+4+4) and the distinct elements evaluated (32 B: 24 B of FP32
    // geometry + 8 B of node ids; the 16 B of node heights in the record
    // are pruning data of this kernel, not algorithmic bytes)
    uint32_t cnt = 0;
    for (uint32_t k = tid; k < new32; k += BS) cnt += __popc(eseen[k]);
    cnt = wave_sum_small<9>(cnt);  // at most 8 words of 32 per thread
    if (lane == 0 && cnt) atomicAdd(&misc[3], cnt);
    // write back decreased copies, lowered interface copies as push records
    // into the item's slot range (entry = interface copies before it);
    // deferred vertices (and sources left by the sweep cap) stay dirty and
    // the submesh is queued again
    uint2 *rec = a.prec + (size_t)inst * a.n_if + si.ioff;
    uint32_t dmin = 0xffffffffu;  // earliest time left dirty, per lane
    for (uint32_t l0 = 0; l0 < nv64; l0 += BS) {
      const uint32_t l = l0 + tid;  // uniform trip count: ballots below
      const uint32_t k = (l0 + wave * 64) >> 5;
      // blocks of 64 copies with nothing lowered or left dirty are skipped
      if (l0 + wave * 64 >= nv64 ||
          (tch[k] | tch[k + 1] | dfr[k] | dfr[k + 1] | (capped ? (flg[k] | flg[k + 1]) : 0u)) == 0u)
        continue;  // uniform across the wave
      bool p = false;
      if (l < nn) {
        if (flag_at(tch, l)) {
          a.phi[cb + l] = phi_l[l];
          p = flag_at(ifc, l);
        }
        if (flag_at(dfr, l) || (capped && flag_at(flg, l))) {
          a.dirty[cb + l] = 1;
          dmin = min(dmin, __float_as_uint(phi_l[l]));
        }
      }
      const uint64_t m = __ballot(p);
      if (m) {
        uint32_t wb = 0;
        if (lane == 0) wb = atomicAdd(&misc[6], (uint32_t)__popcll(m));
        wb = __builtin_amdgcn_readfirstlane(wb);  // every lane of the wave is active
        if (p) {
          const uint32_t w = l >> 5;
          const uint32_t j = ipre[w] + __popc(ifc[w] & ((1u << (l & 31)) - 1u));
          rec[wb + __popcll(m & lanes_below())] = make_uint2(si.ioff + j, __float_as_uint(phi_l[l]));
        }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) dmin = min(dmin, (uint32_t)__shfl_xor(dmin, off));
    if (lane == 0 && dmin != 0xffffffffu) atomicMin(&misc[5], dmin);
    __syncthreads();
    if (tid == 0) {
      push_descriptor(a.cur, cursor.uidx * a.unit + cursor.i, entry, misc[6]);
      abytes += 32ull * misc[3] + 10ull * nn + 12ull * si.ni;
      const float wk = __uint_as_float(misc[5]);
      a.wake[entry] = wk;
      if (wk < F_INF) requeue(a.nxt, a.act, inst, s, a.nsub, a.ninst);
    }
    __syncthreads();
  }
  if (tid == 0 && acts) {
    atomicAdd(&a.stats[0], acts);
    atomicAdd(&a.stats[1], sweeps);
    atomicAdd(&a.stats[2], evals);
    atomicAdd(&a.stats[4], abytes);
  }
  for (int off = 32; off > 0; off >>= 1) vupd += __shfl_xor(vupd, off);
  if (lane == 0 && vupd) atomicAdd(&a.stats[5], vupd);
}


struct SyncArgs {
  const SubInfo *sub;
  const uint4 *if_fo;  // per entry: first other copy, its submesh, then [z, w) in if_os
  const uint2 *if_os;  // per other copy: (copy, its submesh)
  const uint2 *prec;
  float *phi;
  uint8_t *dirty;
  uint32_t *act;
  float *wake;
  WorkList cur, nxt;
  uint32_t *iter;
  uint32_t nsub, ncopy, ninst, n_if, unit, maxch;
};

// interface push: one wave per item slot of the iteration's work units (its
// push descriptor), a lane per push record;
// min over copies, order independent.  The kernel waits on memory, so it
// runs on its own full-occupancy grid.
__global__ __launch_bounds__(256) void k_conv_sync(SyncArgs a) {
  // the compute kernel of this iteration has finished reading the counter
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.iter += 1u;
  const uint32_t lane = threadIdx.x & 63, wpb = blockDim.x >> 6;
  const uint32_t NI = a.cur.cnt[0] * a.unit, nw = gridDim.x * wpb;
  for (uint32_t w = blockIdx.x * wpb + (threadIdx.x >> 6); w < NI; w += nw) {
    const uint2 d = a.cur.desc[w];
    const uint32_t inst = d.x / a.nsub, s = d.x - inst * a.nsub, n = d.y;
    if (n) {
      const uint2 *rec = a.prec + (size_t)inst * a.n_if + a.sub[s].ioff;
      const size_t ib = (size_t)inst * a.ncopy;
      for (uint32_t r = lane; r < n; r += 64) {
        const uint2 e = rec[r];
        const uint32_t v = e.y;
        // most entries have one other copy: the entry's record holds it
        const uint4 fo = a.if_fo[e.x];
        uint2 o = make_uint2(fo.x, fo.y);
        for (uint32_t k = fo.z;;) {
          const size_t pe = (size_t)inst * a.nsub + o.y;
          uint32_t *p = reinterpret_cast<uint32_t *>(a.phi + ib + o.x);
          uint32_t *wk = reinterpret_cast<uint32_t *>(a.wake + pe);
          // Only pushes (no compute kernel runs now) lower phi and wake and
          // only pushes set act, so a stale read errs on the safe side (an
          // extra atomic); the smallest push to a copy always sees v below
          // what it reads, so the flags it sets (dirty, wake, queued) are the
          // same whatever the order, and the minima need no return value.
          // The three reads are issued together.
          const uint32_t cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          const uint32_t wcur = __hip_atomic_load(wk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          const uint32_t queued = __hip_atomic_load(&a.act[pe], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (v < cur) {
            __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            a.dirty[ib + o.x] = 1;
            if (v < wcur) __hip_atomic_fetch_min(wk, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (queued == 0u) enqueue(a.nxt, a.act, inst, o.y, a.nsub, a.ninst);
          }
          if (k >= fo.w) break;
          o = a.if_os[k++];
        }
      }
    }
  }
}

// seeds: (instance, copy) pairs get min(phi, t0), are flagged, and their
// submesh is queued on the current list
__global__ void k_conv_seed(const uint2 *ic, uint32_t n, float t0, const uint32_t *c2s, float *phi,
                            uint8_t *dirty, uint32_t *act, float *wake, WorkList l0, uint32_t nsub,
                            uint32_t ncopy, uint32_t ninst) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t inst = ic[i].x, c = ic[i].y;
    const size_t gc = (size_t)inst * ncopy + c;
    atomicMin(reinterpret_cast<uint32_t *>(phi + gc), __float_as_uint(t0));
    dirty[gc] = 1;
    const uint32_t s = c2s[c];
    atomicMin(reinterpret_cast<uint32_t *>(wake + (size_t)inst * nsub + s), __float_as_uint(t0));
    enqueue(l0, act, inst, s, nsub, ninst);
  }
}

// per instance: 1 if any (instance, submesh) pair is queued
__global__ __launch_bounds__(256) void k_inst_active(const uint32_t *act, uint32_t nsub, uint32_t *out) {
  __shared__ uint32_t any;
  if (threadIdx.x == 0) any = 0u;
  __syncthreads();
  const uint32_t *row = act + (size_t)blockIdx.x * nsub;
  uint32_t v = 0u;
  for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) v |= row[s];
  if (v) atomicOr(&any, 1u);
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = any;
}

// recycles finished instances: times to +inf, flags and wake times cleared,
// and the instance's window restarts at the current iteration
__global__ void k_inst_reset(const uint32_t *insts, uint32_t n, float *phi, uint8_t *dirty, float *wake,
                             uint32_t *start, const uint32_t *iter, uint32_t ncopy, uint32_t nsub) {
  const size_t per = (size_t)ncopy + nsub;
  const size_t total = per * n;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t inst = insts[i / per];
    const size_t k = i % per;
    if (k < ncopy) {
      phi[(size_t)inst * ncopy + k] = F_INF;
      dirty[(size_t)inst * ncopy + k] = 0;
    } else {
      wake[(size_t)inst * nsub + (k - ncopy)] = F_INF;
    }
    if (k == 0) start[inst] = *iter;
  }
}

// warm start: every copy of the instance takes the global field value,
// everything flagged, every submesh queued
__global__ void k_conv_init(const float *phi0, const uint32_t *c2g, uint32_t inst, float *phi,
                            uint8_t *dirty, uint32_t ncopy, int flag) {
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncopy; c += gridDim.x * blockDim.x) {
    const size_t gc = (size_t)inst * ncopy + c;
    phi[gc] = phi0[c2g[c]];
    dirty[gc] = (uint8_t)flag;
  }
}

__global__ void k_conv_queue_all(uint32_t inst, uint32_t nsub, uint32_t ninst, uint32_t *act,
                                 float *wake, WorkList l0) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nsub; s += gridDim.x * blockDim.x) {
    wake[(size_t)inst * nsub + s] = 0.0f;
    enqueue(l0, act, inst, s, nsub, ninst);
  }
}

template <class T>
__global__ void k_fill(T *p, size_t n, T v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// min over copies of each global vertex
__global__ void k_conv_gather(const float *phi, const uint32_t *g2c_dsp, const uint32_t *g2c,
                              uint32_t nn, size_t ib, float *out) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nn; v += gridDim.x * blockDim.x) {
    float m = F_INF;
    for (uint32_t k = g2c_dsp[v]; k < g2c_dsp[v + 1]; ++k) m = fminf(m, phi[ib + g2c[k]]);
    out[v] = m;
  }
}

// residual sweep: max relative amount any copy could still be lowered
__global__ __launch_bounds__(256) void k_conv_residual(const SubInfo *sub, uint32_t nsub,
                                                       const uint2 *e2n, const float *geo,
                                                       const uint8_t *ereg, const float *rscale,
                                                       const float *phi, size_t ib, uint32_t inst,
                                                       uint32_t *res_bits) {
  float worst = 0.0f;
  for (uint32_t s = blockIdx.x; s < nsub; s += gridDim.x) {
    const SubInfo si = sub[s];
    for (uint32_t e = threadIdx.x; e < si.ne; e += blockDim.x) {
      const size_t ge = (size_t)si.eoff + e;
      const uint2 q = e2n[ge];
      const uint32_t nd[4] = {q.x & 0xffffu, q.x >> 16, q.y & 0xffffu, q.y >> 16};
      const float sc = rscale[4 * (size_t)inst + (ereg[ge] & 3)];
      float g[6], ph[4];
      for (int k = 0; k < 6; ++k) g[k] = geo[6 * ge + k] * sc;
      for (int k = 0; k < 4; ++k) ph[k] = phi[ib + si.coff + nd[k]];
      float c[4] = {cand_exact<0, float>(g, ph, F_INF), cand_exact<1, float>(g, ph, F_INF),
                    cand_exact<2, float>(g, ph, F_INF), cand_exact<3, float>(g, ph, F_INF)};
      for (int k = 0; k < 4; ++k)
        if (c[k] < ph[k]) worst = fmaxf(worst, ph[k] == F_INF ? F_INF : (ph[k] - c[k]) / (1.0f + ph[k]));
    }
  }
  if (worst > 0.0f) atomicMax(res_bits, __float_as_uint(worst));
}

// support sweep: for every vertex the smallest candidate its elements give
// (global vertex, float bits, atomicMin); seeds and unreached vertices are
// judged by the caller
__global__ __launch_bounds__(256) void k_conv_support(const SubInfo *sub, uint32_t nsub,
                                                      const uint2 *e2n, const float *geo,
                                                      const uint8_t *ereg, const float *rscale,
                                                      const float *phi, const uint32_t *c2g,
                                                      size_t ib, uint32_t inst, uint32_t *best) {
  for (uint32_t s = blockIdx.x; s < nsub; s += gridDim.x) {
    const SubInfo si = sub[s];
    for (uint32_t e = threadIdx.x; e < si.ne; e += blockDim.x) {
      const size_t ge = (size_t)si.eoff + e;
      const uint2 q = e2n[ge];
      const uint32_t nd[4] = {q.x & 0xffffu, q.x >> 16, q.y & 0xffffu, q.y >> 16};
      const float sc = rscale[4 * (size_t)inst + (ereg[ge] & 3)];
      float g[6], ph[4];
      for (int k = 0; k < 6; ++k) g[k] = geo[6 * ge + k] * sc;
      for (int k = 0; k < 4; ++k) ph[k] = phi[ib + si.coff + nd[k]];
      const float c[4] = {cand_exact<0, float>(g, ph, F_INF), cand_exact<1, float>(g, ph, F_INF),
                          cand_exact<2, float>(g, ph, F_INF), cand_exact<3, float>(g, ph, F_INF)};
      for (int k = 0; k < 4; ++k)
        if (c[k] < F_INF) atomicMin(&best[c2g[si.coff + nd[k]]], __float_as_uint(c[k]));
    }
  }
}

// worst relative amount by which a reached, non-seed vertex lies below the
// best candidate of its elements (a field that is not a fixed point from
// below: an initial field that was not an upper bound)
__global__ void k_support_check(const float *phi_g, const uint32_t *best, uint32_t nn,
                                uint32_t *worst_bits) {
  float worst = 0.0f;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nn; v += gridDim.x * blockDim.x) {
    const float p = phi_g[v];
    if (!(p > 0.0f) || !(p < F_INF)) continue;
    const float b = __uint_as_float(best[v]);
    if (b > p) worst = fmaxf(worst, (b - p) / (1.0f + p));
  }
  if (worst > 0.0f) atomicMax(worst_bits, __float_as_uint(worst));
}

// ---------------------------------------------------------------------------
// compat mode (reference time windows, FP64)
// ---------------------------------------------------------------------------
struct CompatArgs {
  const SubInfo *sub;
  const uint2 *e2n;
  const double *geo;
  const uint32_t *n2e_dsp;
  const uint16_t *n2e_le;
  const uint8_t *n2e_pos;
  const uint32_t *n2n_dsp;
  const uint16_t *n2n_lv;
  double *phi;
  int8_t *st;
  unsigned long long *stats;  // sweeps, slot evaluations
  uint32_t *pending;          // copies left narrow/next after the window
  double *nexte;              // per submesh: earliest window time its copies change
  uint32_t nsub;
  int32_t exact;
  double fw, cur, phi_start, dt, apd;
  uint32_t o_tphi, o_st, o_tst, o_snap, o_misc;
};

__device__ inline uint32_t block_sum(uint32_t v, uint32_t *slot) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (threadIdx.x == 0) *slot = 0u;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) atomicAdd(slot, v);
  __syncthreads();
  const uint32_t r = *slot;
  __syncthreads();
  return r;
}

// FP64 bit pattern -> unsigned key of the same order (no NaN) and back
__device__ inline unsigned long long dkey(double d) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(d);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ inline double dkey_inv(unsigned long long k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & ~(1ull << 63)) : ~k));
}

// minimum over the block, valid in thread 0 only (the slot is reset by
// thread 0 before its first barrier, so back-to-back calls need no third one)
__device__ inline double block_min0(double v, unsigned long long *slot) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
  if (threadIdx.x == 0) *slot = ~0ull;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) atomicMin(slot, dkey(v));
  __syncthreads();
  return threadIdx.x == 0 ? dkey_inv(*slot) : 0.0;
}

// one (vertex, element) slot: src/eikonal/cuda_fim.h:1362-1411.  The slot's
// element nodes nd (local ids) and the owner's position t are given.
__device__ inline double compat_slot(const CompatArgs &a, const SubInfo &si, const double *phi_l,
                                     const int8_t *st_l, size_t ge, const uint32_t nd[4], int t) {
  double ph[4], g[6];
  const double lo = a.phi_start - a.fw;
  for (int j = 0; j < 4; ++j) {
    const uint32_t n = nd[j];
    ph[j] = (j != t && st_l[n] >= FIM_NARROW && phi_l[n] > lo) ? phi_l[n] : DBL_MAX;
  }
  for (int j = 0; j < 6; ++j) g[j] = a.geo[6 * ge + j];
  if (!a.exact) return cand_ref(g, ph, t);
  switch (t) {
    case 0: return cand_exact<0, double>(g, ph, DBL_MAX);
    case 1: return cand_exact<1, double>(g, ph, DBL_MAX);
    case 2: return cand_exact<2, double>(g, ph, DBL_MAX);
    default: return cand_exact<3, double>(g, ph, DBL_MAX);
  }
}

constexpr unsigned long long kDmaxBits = 0x7fefffffffffffffull;  // DBL_MAX

// The candidates of every slot k0 <= k < k1 whose owner has state `want`,
// merged into the owner's minimum (FP64 bits, LDS integer min); returns the
// slots this thread evaluated.
__device__ inline unsigned long long compat_slots(const CompatArgs &a, const SubInfo &si,
                                                  const double *phi_l, const int8_t *st_l, uint32_t k0,
                                                  uint32_t k1, int8_t want, unsigned long long *tmin) {
  unsigned long long n = 0;
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    const int t = a.n2e_pos[k];
    const size_t ge = (size_t)si.eoff + a.n2e_le[k];
    const uint2 q = a.e2n[ge];
    const uint32_t nd[4] = {q.x & 0xffffu, q.x >> 16, q.y & 0xffffu, q.y >> 16};
    const uint32_t owner = nd[t];
    if (st_l[owner] != want) continue;
    const double v = compat_slot(a, si, phi_l, st_l, ge, nd, t);
    atomicMin(&tmin[owner], (unsigned long long)__double_as_longlong(v));
    ++n;
  }
  return n;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_compat_compute(CompatArgs a) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char lds[];
  double *phi_l = reinterpret_cast<double *>(lds);
  double *tphi = reinterpret_cast<double *>(lds + a.o_tphi);
  int8_t *st_l = reinterpret_cast<int8_t *>(lds + a.o_st);
  int8_t *tst = reinterpret_cast<int8_t *>(lds + a.o_tst);
  int8_t *snap = reinterpret_cast<int8_t *>(lds + a.o_snap);
  uint32_t *misc = reinterpret_cast<uint32_t *>(lds + a.o_misc);
  const uint32_t tid = threadIdx.x;
  unsigned long long sweeps = 0, slots = 0;
  unsigned long long *evslot = reinterpret_cast<unsigned long long *>(misc + 2);
  for (uint32_t s = blockIdx.x; s < a.nsub; s += gridDim.x) {
    // a submesh with nothing narrow or next, whose frozen and tail copies all
    // stay put at this window time, would come out of the pass unchanged
    if (a.nexte[s] > a.cur) continue;
    const SubInfo si = a.sub[s];
    const uint32_t nn = si.nn;
    // update_states (cuda_fim.cu:1790-1845), applied to every submesh
    uint32_t cnt = 0;
    double p0 = 0.0;  // the copy as loaded (one copy per thread when nn <= BS)
    int8_t st0 = 0;
    for (uint32_t l = tid; l < nn; l += BS) {
      double p = a.phi[si.coff + l];
      int8_t st = a.st[si.coff + l];
      p0 = p;
      st0 = st;
      if (st == FIM_NARROW) {
        ++cnt;
      } else if (st == FIM_NEXT) {
        ++cnt;
        st = FIM_NARROW;
      } else if (st == FIM_FROZEN) {
        if (!(p + a.apd > a.cur)) st = FIM_TAIL;
      } else if (st == FIM_TAIL) {
        if (!(p + a.apd / 0.91 > a.cur)) {
          p = DBL_MAX;
          st = FIM_FAR;
        }
      }
      phi_l[l] = p;
      st_l[l] = st;
    }
    uint32_t active = block_sum(cnt, misc);
    const bool idle = active == 0;
    // the submesh's (vertex, element) slots, contiguous in the copy-major CSR
    const uint32_t k0 = a.n2e_dsp[si.coff], k1 = a.n2e_dsp[si.coff + nn];
    unsigned long long *tmin = reinterpret_cast<unsigned long long *>(tphi);
    while (active > 0) {
      ++sweeps;
      // a: candidates of narrow vertices, one thread per slot; times are
      // >= 0, so the minimum of the FP64 bit patterns as integers is the
      // minimum of the values and does not depend on the order
      for (uint32_t l = tid; l < nn; l += BS) tmin[l] = kDmaxBits;
      __syncthreads();
      slots += compat_slots(a, si, phi_l, st_l, k0, k1, FIM_NARROW, tmin);
      __syncthreads();
      // b: reduce, window test, convergence test
      for (uint32_t l = tid; l < nn; l += BS) {
        tst[l] = FIM_FAR;
        if (st_l[l] != FIM_NARROW) continue;
        const double old = phi_l[l];
        const double nw = tphi[l] < old ? tphi[l] : old;
        phi_l[l] = nw;
        if (nw - a.phi_start > a.dt)
          st_l[l] = FIM_NEXT;
        else if (fabs(old - nw) < (1e-6 * (1 + (old + nw) * .5)))
          tst[l] = FIM_FROZEN;
      }
      __syncthreads();
      // c: downwind neighbours of just-frozen vertices (flags as after b)
      for (uint32_t l = tid; l < nn; l += BS) snap[l] = tst[l];
      __syncthreads();
      for (uint32_t l = tid; l < nn; l += BS) {
        if (st_l[l] == FIM_NARROW) continue;
        const uint32_t c = si.coff + l;
        for (uint32_t k = a.n2n_dsp[c]; k < a.n2n_dsp[c + 1]; ++k) {
          const uint32_t n = a.n2n_lv[k];
          if (snap[n] == FIM_FROZEN && phi_l[n] < phi_l[l]) {
            tst[l] = st_l[l];
            st_l[l] = FIM_NEIG;
            break;
          }
        }
      }
      __syncthreads();
      // d: candidates of marked vertices, one thread per slot
      for (uint32_t l = tid; l < nn; l += BS) tmin[l] = kDmaxBits;
      __syncthreads();
      slots += compat_slots(a, si, phi_l, st_l, k0, k1, FIM_NEIG, tmin);
      __syncthreads();
      // e: accept improvements
      for (uint32_t l = tid; l < nn; l += BS) {
        if (st_l[l] != FIM_NEIG) continue;
        const double old = phi_l[l];
        const double nw = tphi[l] < old ? tphi[l] : old;
        phi_l[l] = nw;
        st_l[l] = (nw < old) ? (int8_t)FIM_NARROW : tst[l];
      }
      __syncthreads();
      // f: freeze / count
      cnt = 0;
      for (uint32_t l = tid; l < nn; l += BS) {
        if (st_l[l] == FIM_NARROW) {
          if (tst[l] == FIM_FROZEN)
            st_l[l] = FIM_FROZEN;
          else
            ++cnt;
        } else if (st_l[l] == FIM_NEIG) {
          st_l[l] = FIM_NARROW;
          ++cnt;
        }
      }
      active = block_sum(cnt, misc);
    }
    // write back; when nothing was narrow only update_states' time-driven
    // transitions can have changed a copy, and nothing is pending (narrow and
    // next copies are what `active` counted).  ev: the earliest window time
    // at which update_states would change a copy (frozen: p + apd, tail:
    // p + apd / 0.91, as update_states computes them; narrow or next: now)
    const bool lean = idle && nn <= BS;
    uint32_t pend = 0;
    double ev = INFINITY;
    for (uint32_t l = tid; l < nn; l += BS) {
      const double p = phi_l[l];
      const int8_t st = st_l[l];
      if (!lean || p != p0 || st != st0) {
        a.phi[si.coff + l] = p;
        a.st[si.coff + l] = st;
      }
      if (st == FIM_NARROW || st == FIM_NEXT) {
        ++pend;
        ev = -INFINITY;
      } else if (st == FIM_FROZEN) {
        ev = fmin(ev, p + a.apd);
      } else if (st == FIM_TAIL) {
        ev = fmin(ev, p + a.apd / 0.91);
      }
    }
    if (!lean) {
      pend = block_sum(pend, misc);
      if (tid == 0 && pend) atomicAdd(a.pending, pend);
    }
    ev = block_min0(ev, evslot);
    if (tid == 0) a.nexte[s] = ev;
  }
  if (tid == 0 && sweeps) atomicAdd(&a.stats[0], sweeps);
  for (int off = 32; off > 0; off >>= 1) slots += __shfl_xor(slots, off);
  if ((tid & 63) == 0 && slots) atomicAdd(&a.stats[1], slots);
}

// sync_step2 over vertices with several copies
__global__ void k_compat_sync(const uint32_t *grp_v, uint32_t ngrp, const uint32_t *g2c_dsp,
                              const uint32_t *g2c, const uint32_t *c2s, double *phi, int8_t *st,
                              double *nexte) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ngrp; i += gridDim.x * blockDim.x) {
    const uint32_t v = grp_v[i];
    double mn = DBL_MAX;
    for (uint32_t k = g2c_dsp[v]; k < g2c_dsp[v + 1]; ++k) {
      const double c = phi[g2c[k]];
      if (c < mn) mn = c;
    }
    for (uint32_t k = g2c_dsp[v]; k < g2c_dsp[v + 1]; ++k) {
      const uint32_t c = g2c[k];
      if (mn < phi[c]) {
        phi[c] = mn;
        st[c] = FIM_NARROW;
        nexte[c2s[c]] = -INFINITY;
      }
    }
  }
}

__global__ void k_compat_seed(const uint2 *ic, uint32_t n, double t0, const uint32_t *c2s, double *phi,
                              int8_t *st, double *nexte) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    phi[ic[i].y] = t0;
    st[ic[i].y] = FIM_NARROW;
    nexte[c2s[ic[i].y]] = -INFINITY;
  }
}

__global__ void k_compat_gather(const double *phi, const int8_t *st, const uint32_t *g2c_dsp,
                                const uint32_t *g2c, uint32_t nn, double *out_phi, int32_t *out_st) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nn; v += gridDim.x * blockDim.x) {
    double m = DBL_MAX;
    int32_t s = FIM_FAR;
    for (uint32_t k = g2c_dsp[v]; k < g2c_dsp[v + 1]; ++k) {
      const uint32_t c = g2c[k];
      if (phi[c] < m) m = phi[c];
      if (st[c] > s) s = st[c];
    }
    out_phi[v] = m;
    out_st[v] = s;
  }
}

uint32_t al16(uint32_t x) { return (x + 15u) & ~15u; }

template <class T>
T *dalloc(size_t n, std::vector<void *> &owned) {
  void *p = nullptr;
  if (n == 0) n = 1;
  HIPCHECK(hipMalloc(&p, n * sizeof(T)));
  owned.push_back(p);
  return static_cast<T *>(p);
}

template <class T>
T *dupload(const std::vector<T> &v, std::vector<void *> &owned) {
  T *p = dalloc<T>(v.size(), owned);
  if (!v.empty()) HIPCHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

}  // namespace
}  // namespace fim

// Scratch of fim_support_noalloc: the best candidate time (float bits) of
// every vertex, for meshes of up to kNoAllocMaxVertices vertices (4 MiB).
// One copy per device, held until the process ends.  Host code never uses
// this name as an address: fim_create looks the device address up once per
// device (noalloc_scratch) and the kernels get that pointer.
constexpr uint32_t kNoAllocMaxVertices = 1u << 20;
static __device__ __align__(256) uint32_t scratch[kNoAllocMaxVertices];

using namespace fim;

struct fim_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  fim_opts o{};
  uint32_t nsub = 0, ncopy = 0, ninst = 1, nn = 0;
  int64_t ne = 0;
  int32_t nn_max = 0, ne_max = 0;
  std::vector<void *> owned;
  // host copies for seeding
  std::vector<uint32_t> g2c_dsp, g2c;
  // pack
  SubInfo *d_sub = nullptr;
  uint2 *d_e2n = nullptr;
  float *d_geo32 = nullptr;
  float4 *d_hgt = nullptr;
  double *d_geo64 = nullptr;
  uint8_t *d_ereg = nullptr;
  uint32_t *d_c2g = nullptr, *d_c2s = nullptr, *d_g2c_dsp = nullptr, *d_g2c = nullptr, *d_grp = nullptr;
  uint32_t ngrp = 0;
  uint16_t *d_if_l = nullptr;
  uint4 *d_if_fo = nullptr;
  uint2 *d_if_os = nullptr;
  uint32_t *d_n2e_dsp = nullptr, *d_n2n_dsp = nullptr;
  uint16_t *d_n2e_le = nullptr, *d_n2n_lv = nullptr;
  uint8_t *d_n2e_pos = nullptr;
  uint4 *d_erec = nullptr, *d_adj = nullptr;
  uint32_t *d_adj_ptr = nullptr;
  // converge state
  float *d_phi = nullptr, *d_rscale = nullptr;
  uint8_t *d_dirty = nullptr;
  uint32_t *d_act = nullptr, *d_cnt = nullptr;
  uint32_t *d_start = nullptr;  // n_inst: iteration each instance was (re)seeded at
  float *d_winst = nullptr;     // n_inst: time window per iteration (us), 0 = off
  // per-instance region factors and windows as set on the host
  // (fim_set_scaling, fim_set_window), uploaded together before the next
  // launch that reads them: setting up a batch costs no device round trip
  // per instance
  std::vector<float> h_rscale, h_winst;
  bool params_dirty = false;
  uint32_t *d_flag = nullptr;   // n_inst: scratch for instance queries / lists
  uint2 *d_prec = nullptr;  // push records: n_inst x (interface entries)
  uint2 *d_ifm = nullptr;
  uint32_t wps = 1;
  WorkList wl[3] = {};
  uint32_t unit = 2, maxch = 1;
  float *d_wake = nullptr;
  // compat state
  double *d_phi64 = nullptr;
  int8_t *d_st = nullptr;
  double *d_nexte = nullptr;  // per submesh: earliest window time a copy changes
  uint32_t *d_pending = nullptr;
  // shared
  unsigned long long *d_stats = nullptr;
  uint32_t *h_pin = nullptr;
  // seed staging: pinned host and device buffers of (instance, copy) pairs
  uint2 *h_seed = nullptr, *d_seed = nullptr;
  size_t seed_cap = 0;
  float *d_tmpf = nullptr;
  double *d_tmpd = nullptr;
  int32_t *d_tmpi = nullptr;
  uint32_t *d_tmpu = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evc = nullptr;  // end of an iteration chunk's read-back
  std::vector<hipEvent_t> kev;  // profile: events around each captured compute launch
  ConvArgs ca{};
  SyncArgs sa{};
  CompatArgs xa{};
  uint32_t lds_conv = 0, lds_compat = 0;
  int grid = 0, sync_grid = 0, block = 256, graph_iters = 8;
  int band_wpe = 8;  // waves per SIMD the band kernel's registers are sized for
  bool band_pairs = false;  // band kernel evaluates (element, target) pairs
  bool direct = false;      // launch iterations directly instead of replaying the graph
  fim_stats_t stats{};
};

namespace {

const void *conv_kernel_ptr(bool staged, int block, int wpe, bool pairs);

void launch_conv_compute(fim_ctx *c, const ConvArgs &a) {
  if (c->o.staged) {
    switch (c->block) {
      case 64: k_conv_compute<64, true><<<c->grid, 64, c->lds_conv, c->stream>>>(a); break;
      case 128: k_conv_compute<128, true><<<c->grid, 128, c->lds_conv, c->stream>>>(a); break;
      case 512: k_conv_compute<512, true><<<c->grid, 512, c->lds_conv, c->stream>>>(a); break;
      default: k_conv_compute<256, true><<<c->grid, 256, c->lds_conv, c->stream>>>(a); break;
    }
  } else {
    switch (c->block) {
      case 64: k_conv_band<64, 8, false><<<c->grid, 64, c->lds_conv, c->stream>>>(a); break;
      case 128: k_conv_band<128, 8, false><<<c->grid, 128, c->lds_conv, c->stream>>>(a); break;
      case 512:
        if (c->band_pairs) k_conv_band<512, 8, true><<<c->grid, 512, c->lds_conv, c->stream>>>(a);
        else k_conv_band<512, 8, false><<<c->grid, 512, c->lds_conv, c->stream>>>(a);
        break;
      default: {
        const void *f = conv_kernel_ptr(false, 256, c->band_wpe, c->band_pairs);
        void *args[] = {const_cast<ConvArgs *>(&a)};
        HIPCHECK(hipLaunchKernel(f, dim3(c->grid), dim3(256), args, c->lds_conv, c->stream));
        break;
      }
    }
  }
  HIPCHECK(hipGetLastError());
}

const void *conv_kernel_ptr(bool staged, int block, int wpe, bool pairs) {
  switch (block) {
    case 64: return staged ? (const void *)k_conv_compute<64, true> : (const void *)k_conv_band<64, 8, false>;
    case 128: return staged ? (const void *)k_conv_compute<128, true> : (const void *)k_conv_band<128, 8, false>;
    case 512: return staged ? (const void *)k_conv_compute<512, true>
                            : (pairs ? (const void *)k_conv_band<512, 8, true> : (const void *)k_conv_band<512, 8, false>);
    default:
      if (staged) return (const void *)k_conv_compute<256, true>;
      if (pairs) {
        if (wpe == 5) return (const void *)k_conv_band<256, 5, true>;
        if (wpe == 6) return (const void *)k_conv_band<256, 6, true>;
        if (wpe == 7) return (const void *)k_conv_band<256, 7, true>;
        return (const void *)k_conv_band<256, 8, true>;
      }
      if (wpe == 5) return (const void *)k_conv_band<256, 5, false>;
      if (wpe == 6) return (const void *)k_conv_band<256, 6, false>;
      if (wpe == 7) return (const void *)k_conv_band<256, 7, false>;
      return (const void *)k_conv_band<256, 8, false>;
  }
}

void set_lds_attr(fim_ctx *c) {
  const void *f = conv_kernel_ptr(c->o.staged != 0, c->block, c->band_wpe, c->band_pairs);
  HIPCHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds_conv));
  const void *g = nullptr;
  switch (c->block) {
    case 64: g = (const void *)k_compat_compute<64>; break;
    case 128: g = (const void *)k_compat_compute<128>; break;
    case 512: g = (const void *)k_compat_compute<512>; break;
    default: g = (const void *)k_compat_compute<256>; break;
  }
  HIPCHECK(hipFuncSetAttribute(g, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds_compat));
}

void launch_compat_compute(fim_ctx *c, const CompatArgs &a) {
  // 8192 workgroups: more submeshes in flight than 4096 (-4 % per window on
  // the slab's 10,162), no further gain from one workgroup per submesh
  const int grid = (int)std::min<uint32_t>(c->nsub, 8192u);
  switch (c->block) {
    case 64: k_compat_compute<64><<<grid, 64, c->lds_compat, c->stream>>>(a); break;
    case 128: k_compat_compute<128><<<grid, 128, c->lds_compat, c->stream>>>(a); break;
    case 512: k_compat_compute<512><<<grid, 512, c->lds_compat, c->stream>>>(a); break;
    default: k_compat_compute<256><<<grid, 256, c->lds_compat, c->stream>>>(a); break;
  }
  HIPCHECK(hipGetLastError());
}

int blocks_for(size_t n, int bs) { return (int)std::max<size_t>(1, std::min<size_t>((n + bs - 1) / bs, 65536)); }

// work units of list l; the buckets of list z (consumed), if any, are cleared
void launch_plan(fim_ctx *c, int l, int z) {
  PlanArgs pa{c->wl[l].bcnt, c->wl[l].plan, c->wl[l].cnt, z >= 0 ? c->wl[z].bcnt : nullptr,
              c->nsub, c->unit, c->maxch};
  k_conv_plan<<<1, 1024, 0, c->stream>>>(pa);
  HIPCHECK(hipGetLastError());
}

// one global iteration k: work lists rotate over three buffers; iteration k
// reads list k%3 and appends to (k+1)%3, whose units are planned after the
// interface push; the buckets of list k%3 are cleared then
void enqueue_iteration(fim_ctx *c, int k, bool timed) {
  const int p = k % 3, q = (k + 1) % 3;
  if (timed) HIPCHECK(hipEventRecord(c->kev[2 * k], c->stream));
  ConvArgs a = c->ca;
  a.cur = c->wl[p];
  a.nxt = c->wl[q];
  launch_conv_compute(c, a);
  if (timed) HIPCHECK(hipEventRecord(c->kev[2 * k + 1], c->stream));
  SyncArgs s = c->sa;
  s.cur = c->wl[p];
  s.nxt = c->wl[q];
  k_conv_sync<<<c->sync_grid, 256, 0, c->stream>>>(s);
  HIPCHECK(hipGetLastError());
  launch_plan(c, q, p);
}

void build_graph(fim_ctx *c) {
  HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int i = 0; i < c->graph_iters; ++i) enqueue_iteration(c, i, false);
  } catch (...) {
    hipGraph_t g;
    (void)hipStreamEndCapture(c->stream, &g);
    throw;
  }
  HIPCHECK(hipStreamEndCapture(c->stream, &c->graph));
  HIPCHECK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
}

void check_inst(const fim_ctx *c, int32_t inst) {
  if (inst < 0 || (uint32_t)inst >= c->ninst) throw Error(FIM_EINVAL, "instance index out of range");
}

std::vector<uint32_t> copies_of(const fim_ctx *c, const int64_t *vtx, int64_t n) {
  std::vector<uint32_t> out;
  for (int64_t i = 0; i < n; ++i) {
    if (vtx[i] < 0 || vtx[i] >= (int64_t)c->nn) throw Error(FIM_EINVAL, "seed vertex out of range");
    for (uint32_t k = c->g2c_dsp[vtx[i]]; k < c->g2c_dsp[vtx[i] + 1]; ++k) out.push_back(c->g2c[k]);
  }
  return out;
}

void destroy_ctx(fim_ctx *c) {
  if (!c) return;
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->evc) (void)hipEventDestroy(c->evc);
  for (hipEvent_t e : c->kev)
    if (e) (void)hipEventDestroy(e);
  for (void *p : c->owned) (void)hipFree(p);
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  if (c->h_seed) (void)hipHostFree(c->h_seed);
  if (c->d_seed) (void)hipFree(c->d_seed);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

template <class F>
int guarded(F &&f) {
  try {
    f();
    return FIM_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return FIM_EINVAL;
  } catch (...) {
    set_last_error("unknown error");
    return FIM_EINVAL;
  }
}

}  // namespace

namespace fim {

void *ctx_stream(fim_ctx *c) { return c->stream; }
int32_t ctx_device(const fim_ctx *c) { return c->dev; }
uint32_t ctx_vertices(const fim_ctx *c) { return c->nn; }
int32_t ctx_instances(const fim_ctx *c) { return (int32_t)c->ninst; }
bool ctx_converge(const fim_ctx *c) { return c->o.mode == FIM_MODE_CONVERGE; }

const float *ctx_gather(fim_ctx *c, int32_t inst) {
  check_inst(c, inst);
  k_conv_gather<<<blocks_for(c->nn, 256), 256, 0, c->stream>>>(c->d_phi, c->d_g2c_dsp, c->d_g2c, c->nn,
                                                               (size_t)inst * c->ncopy, c->d_tmpf);
  HIPCHECK(hipGetLastError());
  return c->d_tmpf;
}

float *ctx_field_buffer(fim_ctx *c) { return c->d_tmpf; }

void ctx_init_field(fim_ctx *c, int32_t inst, bool all_active) {
  check_inst(c, inst);
  k_conv_init<<<blocks_for(c->ncopy, 256), 256, 0, c->stream>>>(c->d_tmpf, c->d_c2g, (uint32_t)inst,
                                                                 c->d_phi, c->d_dirty, c->ncopy,
                                                                 all_active ? 1 : 0);
  if (all_active)
    k_conv_queue_all<<<blocks_for(c->nsub, 256), 256, 0, c->stream>>>((uint32_t)inst, c->nsub, c->ninst,
                                                                       c->d_act, c->d_wake, c->wl[0]);
  HIPCHECK(hipGetLastError());
}

// device address of `scratch` on the current device `dev`, looked up once
// per device (hipGetSymbolAddress) and kept
uint32_t *noalloc_scratch(int dev) {
  static std::mutex mu;
  static std::vector<uint32_t *> ptr;
  std::lock_guard<std::mutex> lock(mu);
  if (dev < 0) throw Error(FIM_EINVAL, "bad device ordinal");
  if ((size_t)dev >= ptr.size()) ptr.resize((size_t)dev + 1, nullptr);
  if (!ptr[dev]) {
    void *p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(FIM_EDEVICE, "no device address for the support scratch");
    ptr[dev] = static_cast<uint32_t *>(p);
  }
  return ptr[dev];
}

// uploads the per-instance factors and windows set since the last upload
// (the device must be current)
void flush_params(fim_ctx *c) {
  if (!c->params_dirty) return;
  HIPCHECK(hipMemcpyAsync(c->d_rscale, c->h_rscale.data(), c->h_rscale.size() * sizeof(float),
                          hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->d_winst, c->h_winst.data(), c->h_winst.size() * sizeof(float),
                          hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->params_dirty = false;
}

// support check of one instance with `best` (nn words of device memory) as
// the per-vertex candidate minimum
double support_into(fim_ctx *c, int32_t inst, uint32_t *best) {
  flush_params(c);
  k_fill<uint32_t><<<blocks_for(c->nn, 256), 256, 0, c->stream>>>(best, c->nn, 0x7f800000u);
  k_conv_support<<<(int)std::min<uint32_t>(c->nsub, 4096u), 256, 0, c->stream>>>(
      c->d_sub, c->nsub, c->d_e2n, c->d_geo32, c->d_ereg, c->d_rscale, c->d_phi, c->d_c2g,
      (size_t)inst * c->ncopy, (uint32_t)inst, best);
  k_conv_gather<<<blocks_for(c->nn, 256), 256, 0, c->stream>>>(c->d_phi, c->d_g2c_dsp, c->d_g2c, c->nn,
                                                               (size_t)inst * c->ncopy, c->d_tmpf);
  k_fill<uint32_t><<<1, 64, 0, c->stream>>>(c->d_tmpu, 1, 0u);
  k_support_check<<<blocks_for(c->nn, 256), 256, 0, c->stream>>>(c->d_tmpf, best, c->nn, c->d_tmpu);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(c->h_pin, c->d_tmpu, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  float r;
  std::memcpy(&r, c->h_pin, 4);
  return r;
}

}  // namespace fim

extern "C" {

int fim_create(const fim_pack *pk, const fim_opts *o, fim_ctx **out) {
  return guarded([&] {
    if (!pk || !o || !out) throw Error(FIM_EINVAL, "null argument");
    *out = nullptr;
    const Pack &p = pk->p;
    fim_ctx *c = new fim_ctx();
    try {
      c->o = *o;
      c->dev = o->device;
      HIPCHECK(hipSetDevice(c->dev));
      (void)noalloc_scratch(c->dev);  // once per device: fim_support_noalloc's scratch
      // CUFIM_DIRECT_LAUNCH=1 launches the iterations' kernels directly instead
      // of replaying a captured graph (the same kernel sequence): for tools
      // that cannot trace graph launches
      {
        const char *dl = std::getenv("CUFIM_DIRECT_LAUNCH");
        c->direct = dl && dl[0] == '1';
      }
      HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      HIPCHECK(hipEventCreate(&c->ev0));
      HIPCHECK(hipEventCreate(&c->ev1));
      HIPCHECK(hipEventCreateWithFlags(&c->evc, hipEventDisableTiming));
      if (o->mode != FIM_MODE_CONVERGE && o->mode != FIM_MODE_COMPAT)
        throw Error(FIM_EINVAL, "unknown mode");
      c->block = o->block ? o->block : (o->mode == FIM_MODE_CONVERGE ? 512 : 256);
      if (c->block != 64 && c->block != 128 && c->block != 256 && c->block != 512)
        throw Error(FIM_EINVAL, "block must be 64, 128, 256 or 512");
      if (o->waves < 0 || o->waves > 8) throw Error(FIM_EINVAL, "waves must be 0..8");
      c->ninst = o->mode == FIM_MODE_CONVERGE ? (uint32_t)std::max(1, o->n_instances) : 1u;
      // a multiple of 3 so every graph launch starts on list 0
      c->graph_iters = o->graph_iters > 0 ? 3 * ((o->graph_iters + 2) / 3) : 6;
      c->nsub = (uint32_t)p.nsub;
      c->ncopy = (uint32_t)p.ncopy;
      c->nn = (uint32_t)p.nn_global;
      c->ne = p.ne_global;
      c->nn_max = p.nn_max;
      c->ne_max = p.ne_max;
      if ((uint64_t)c->ninst * c->nsub >= (1ull << 32) || (uint64_t)c->ninst * c->ncopy >= (1ull << 40))
        throw Error(FIM_ELIMIT, "too many instances for this mesh");
      c->g2c_dsp = p.g2c_dsp;
      c->g2c = p.g2c;
      auto &ow = c->owned;
      c->d_sub = dupload(p.sub, ow);
      {
        std::vector<uint2> e2(p.ne_global);
        for (int64_t e = 0; e < p.ne_global; ++e) {
          e2[e].x = (uint32_t)p.e2n_l[4 * e] | ((uint32_t)p.e2n_l[4 * e + 1] << 16);
          e2[e].y = (uint32_t)p.e2n_l[4 * e + 2] | ((uint32_t)p.e2n_l[4 * e + 3] << 16);
        }
        c->d_e2n = dupload(e2, ow);
      }
      c->d_ereg = dupload(p.ereg, ow);
      c->d_c2g = dupload(p.c2g, ow);
      c->d_c2s = dupload(p.c2s, ow);
      c->d_g2c_dsp = dupload(p.g2c_dsp, ow);
      c->d_g2c = dupload(p.g2c, ow);
      c->d_stats = dalloc<unsigned long long>(8, ow);
      HIPCHECK(hipHostMalloc((void **)&c->h_pin, 64, hipHostMallocDefault));
      c->d_tmpu = dalloc<uint32_t>(4, ow);
      HIPCHECK(hipMemset(c->d_stats, 0, 8 * sizeof(unsigned long long)));
      if (o->mode == FIM_MODE_CONVERGE) {
        c->lds_conv = 0;
        const uint32_t nn64 = ((uint32_t)c->nn_max + 63u) & ~63u;
        const uint32_t ne32 = ((uint32_t)c->ne_max + 31u) / 32u;
        const uint32_t o_phin = al16(4 * nn64);
        const uint32_t o_flg = o_phin + al16(4 * nn64);
        const uint32_t o_tch = o_flg + al16(4 * (nn64 / 32));
        const uint32_t o_dfr = o_tch + al16(4 * (nn64 / 32));
        const uint32_t o_e2n = o_dfr + al16(4 * (nn64 / 32));
        // staged: element nodes + geometry; unstaged: element mark bitmasks
        const uint32_t o_geo = o_e2n + (o->staged ? al16(8 * (uint32_t)c->ne_max) : al16(4 * ne32));
        const uint32_t o_list = o_geo + (o->staged ? al16(24 * (uint32_t)c->ne_max) : al16(4 * ne32));
        // workgroups of 256 and 512 evaluate (element, target) pairs when
        // local element ids fit 14 bits; the pair rings live in the source list
        c->band_pairs = !o->staged && (c->block == 256 || c->block == 512) && c->ne_max < (1 << 14);
        const uint32_t vlist_bytes =
            o->staged ? 0u
                      : al16(std::max<uint32_t>(2 * nn64, c->band_pairs ? (uint32_t)(c->block / 64) * 2u * kPairRing : 0u));
        const uint32_t tail_bytes = vlist_bytes + 48 + (o->staged ? 0u : 3u * al16(4 * (nn64 / 32)));
        // The band kernel lists at most elcap elements per pass: at least
        // kElistCap, more if the LDS left over by the workgroups the wave
        // slots admit (8 waves per SIMD) holds more
        uint32_t elcap = (uint32_t)c->ne_max;
        if (!o->staged) {
          const uint32_t wg_slots = std::max<uint32_t>(1u, 32u / ((uint32_t)c->block / 64u));
          const uint32_t budget = (160u * 1024u / wg_slots) & ~15u;
          const uint32_t base = o_list + tail_bytes;
          const uint32_t room = budget > base ? ((budget - base) / 2u) & ~7u : 0u;
          elcap = std::min<uint32_t>((uint32_t)c->ne_max, std::max<uint32_t>((uint32_t)kElistCap, room));
        }
        const uint32_t o_vlist = o_list + al16(2 * elcap);
        const uint32_t o_misc = o_vlist + vlist_bytes;
        const uint32_t o_wrt = o_misc + 48;  // misc: 8 words, 4 region factors
        // unstaged: interface-copy bits and entry counts (2 x nn64/32 words)
        const uint32_t o_ifc = o_wrt + (o->staged ? 0u : al16(4 * (nn64 / 32)));
        c->lds_conv = o_ifc + (o->staged ? 0u : 2u * al16(4 * (nn64 / 32)));
        if (c->lds_conv > 160 * 1024)
          throw Error(FIM_ELIMIT, "largest submesh needs " + std::to_string(c->lds_conv) +
                                      " B of LDS (limit 163840); use a smaller partition budget");
        c->d_geo32 = dupload(p.geo32, ow);
        {
          std::vector<float4> hg(p.ne_global);
          for (int64_t e = 0; e < p.ne_global; ++e)
            hg[e] = make_float4(p.hgt32[4 * e], p.hgt32[4 * e + 1], p.hgt32[4 * e + 2], p.hgt32[4 * e + 3]);
          c->d_hgt = dupload(hg, ow);
        }
        c->d_n2e_dsp = dupload(p.n2e_dsp, ow);
        c->d_n2e_le = dupload(p.n2e_le, ow);
        if (!o->staged) {
          if (c->nn_max > 32768) throw Error(FIM_ELIMIT, "submeshes above 32768 copies");
          std::vector<uint4> rec(3 * (size_t)p.ne_global);
          for (int64_t e = 0; e < p.ne_global; ++e) {
            const uint16_t *n = &p.e2n_l[4 * e];
            const uint32_t r = p.ereg.empty() ? 0u : (p.ereg[e] & 3u);
            const float *g = &p.geo32[6 * e], *h = &p.hgt32[4 * e];
            auto fb = [](float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; };
            rec[3 * e] = make_uint4((uint32_t)n[0] | ((uint32_t)n[1] << 16) | ((r & 1u) << 15) |
                                        ((r & 2u) << 30),
                                    (uint32_t)n[2] | ((uint32_t)n[3] << 16), fb(g[0]), fb(g[1]));
            rec[3 * e + 1] = make_uint4(fb(g[2]), fb(g[3]), fb(g[4]), fb(g[5]));
            rec[3 * e + 2] = make_uint4(fb(h[0]), fb(h[1]), fb(h[2]), fb(h[3]));
          }
          c->d_erec = dupload(rec, ow);
          std::vector<uint32_t> aptr(p.ncopy);
          std::vector<uint16_t> adj;
          for (int64_t cc = 0; cc < p.ncopy; ++cc) {
            const uint32_t kb = p.n2e_dsp[cc], ke = p.n2e_dsp[cc + 1];
            const uint32_t nch = (ke - kb + 7) / 8;
            if (nch == 0 || nch > 63) throw Error(FIM_ELIMIT, "vertex with no or over 504 incident elements");
            const size_t off = adj.size() / 8;
            if (off >= (1u << 26)) throw Error(FIM_ELIMIT, "adjacency above 2^26 chunks");
            aptr[cc] = (uint32_t)(off << 6) | nch;
            for (uint32_t k = kb; k < ke; ++k) adj.push_back(p.n2e_le[k]);
            while (adj.size() % 8) adj.push_back(0xffffu);
          }
          c->d_adj_ptr = dupload(aptr, ow);
          std::vector<uint4> adj4(adj.size() / 8);
          std::memcpy(adj4.data(), adj.data(), adj.size() * 2);
          c->d_adj = dupload(adj4, ow);
        }
        c->d_if_l = dupload(p.if_l, ow);
        {
          std::vector<uint2> os(p.if_other.size());
          for (size_t k = 0; k < os.size(); ++k) os[k] = make_uint2(p.if_other[k], p.c2s[p.if_other[k]]);
          std::vector<uint4> fo(p.if_l.size());
          for (size_t e = 0; e < fo.size(); ++e) {
            const uint32_t k0 = p.if_dsp[e], k1 = p.if_dsp[e + 1];
            if (k1 <= k0) throw Error(FIM_EINVAL, "interface entry without another copy");
            fo[e] = make_uint4(os[k0].x, os[k0].y, k0 + 1, k1);
          }
          c->d_if_os = dupload(os, ow);
          c->d_if_fo = dupload(fo, ow);
        }
        const size_t ncs = (size_t)c->ninst * c->ncopy, nss = (size_t)c->ninst * c->nsub;
        // push descriptors name an item by instance * nsub + submesh in 32 bits
        if (nss >= (1ull << 32)) throw Error(FIM_ELIMIT, "instances x submeshes exceed 2^32");
        c->d_phi = dalloc<float>(ncs, ow);
        c->d_dirty = dalloc<uint8_t>(ncs, ow);
        {
          // push records: one slot per (instance, interface entry)
          c->d_prec = dalloc<uint2>((size_t)c->ninst * p.if_l.size(), ow);
          c->d_start = dalloc<uint32_t>(c->ninst, ow);
          c->h_winst.assign(c->ninst, (float)std::max(0.0, o->window));
          c->d_winst = dupload(c->h_winst, ow);
          c->d_flag = dalloc<uint32_t>(c->ninst, ow);
          HIPCHECK(hipMemset(c->d_start, 0, sizeof(uint32_t) * c->ninst));
          // per submesh: interface-copy bits of each 32 local copies and the
          // number of interface entries before that word (if_l ascends)
          c->wps = (uint32_t)((c->nn_max + 63) / 64 * 2);
          std::vector<uint2> ifm((size_t)c->nsub * c->wps, make_uint2(0u, 0u));
          for (uint32_t s = 0; s < c->nsub; ++s) {
            const SubInfo &si = p.sub[s];
            for (uint32_t j = 0; j < si.ni; ++j) {
              const uint32_t l = p.if_l[si.ioff + j];
              if (j > 0 && l <= p.if_l[si.ioff + j - 1]) throw Error(FIM_EINVAL, "interface list not ascending");
              ifm[(size_t)s * c->wps + (l >> 5)].x |= 1u << (l & 31);
            }
            uint32_t run = 0;
            for (uint32_t w = 0; w < c->wps; ++w) {
              ifm[(size_t)s * c->wps + w].y = run;
              run += (uint32_t)__builtin_popcount(ifm[(size_t)s * c->wps + w].x);
            }
          }
          c->d_ifm = dupload(ifm, ow);
        }
        c->d_act = dalloc<uint32_t>(nss, ow);
        c->d_wake = dalloc<float>(nss, ow);
        c->unit = o->unit > 0 ? (uint32_t)o->unit : 2u;
        c->maxch = (c->ninst + c->unit - 1) / c->unit;
        if ((uint64_t)c->nsub * c->maxch >= (1ull << 32)) throw Error(FIM_ELIMIT, "too many work units");
        // 16 words per list (units, pairs, 8 XCD counters), then the
        // iteration counter at word 48
        c->d_cnt = dalloc<uint32_t>(64, ow);
        for (int k = 0; k < 3; ++k) {
          c->wl[k].bcnt = dalloc<uint32_t>(c->nsub, ow);
          c->wl[k].bins = dalloc<uint32_t>(nss, ow);
          c->wl[k].plan = dalloc<uint32_t>((size_t)c->nsub * c->maxch, ow);
          c->wl[k].cnt = c->d_cnt + 16 * k;
          c->wl[k].desc = dalloc<uint2>((size_t)c->nsub * c->maxch * c->unit, ow);
        }
        c->d_rscale = dalloc<float>(4 * (size_t)c->ninst, ow);
        c->d_tmpf = dalloc<float>(c->nn, ow);
        k_fill<float><<<blocks_for(4 * (size_t)c->ninst, 256), 256, 0, c->stream>>>(c->d_rscale, 4 * (size_t)c->ninst, 1.0f);
        c->h_rscale.assign(4 * (size_t)c->ninst, 1.0f);
        int ncu = 256;
        HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev));
        // one resident wave of workgroups: registers, LDS and wave slots.
        // Workgroups of 256 put one wave on each SIMD, so the LDS carve
        // alone caps the waves per SIMD at 160 KB / carve; the band kernel
        // variant sized for that many waves gets the registers left over.
        {
          const uint32_t lds_cap = 160u * 1024u / std::max<uint32_t>(c->lds_conv, 1u);
          c->band_wpe = lds_cap <= 5u ? 5 : (lds_cap >= 8u ? 8 : (int)lds_cap);
          if (o->waves > 0) c->band_wpe = std::min(c->band_wpe, std::max(o->waves, 5));
        }
        set_lds_attr(c);
        int per_cu = 0;
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, conv_kernel_ptr(o->staged != 0, c->block, c->band_wpe, c->band_pairs), c->block,
            c->lds_conv));
        per_cu = std::max(per_cu, 1);
        c->grid = o->grid > 0 ? o->grid : ncu * per_cu;
        c->grid = std::max(c->grid, (int)kXcd);  // every XCD eighth needs a workgroup
        c->sync_grid = ncu * 8;  // eight workgroups of 256 fill a CU's wave slots
        c->lds_compat = 0;
        c->ca = ConvArgs{c->d_sub, c->d_e2n, c->d_geo32, c->d_hgt, c->d_ereg, c->d_rscale, c->d_phi,
                         c->d_dirty, c->d_act, c->d_wake, c->d_if_l, c->d_ifm, c->d_prec,
                         WorkList{}, WorkList{}, c->d_stats, c->d_cnt + 48, c->d_start,
                         c->d_n2e_dsp, c->d_n2e_le, c->d_erec, c->d_adj_ptr, c->d_adj,
                         c->nsub, c->ncopy, c->ninst, (float)(o->eps_rel > 0 ? o->eps_rel : 1e-6),
                         c->d_winst, o->local_iters, c->unit, c->maxch, (uint32_t)p.if_l.size(),
                         c->wps, elcap, o_phin, o_flg, o_tch, o_dfr, o_e2n, o_geo, o_list, o_vlist, o_misc,
                         o_wrt, o_ifc};
        c->sa = SyncArgs{c->d_sub, c->d_if_fo, c->d_if_os, c->d_prec,
                         c->d_phi, c->d_dirty, c->d_act, c->d_wake, WorkList{}, WorkList{},
                         c->d_cnt + 48, c->nsub, c->ncopy, c->ninst, (uint32_t)p.if_l.size(),
                         c->unit, c->maxch};
      } else {
        const uint32_t nn = (uint32_t)c->nn_max;
        const uint32_t o_tphi = al16(8 * nn);
        const uint32_t o_st = o_tphi + al16(8 * nn);
        const uint32_t o_tst = o_st + al16(nn);
        const uint32_t o_snap = o_tst + al16(nn);
        const uint32_t o_misc = o_snap + al16(nn);
        c->lds_compat = o_misc + 16;
        if (c->lds_compat > 160 * 1024)
          throw Error(FIM_ELIMIT, "largest submesh too large for compat mode LDS");
        c->d_geo64 = dupload(p.geo64, ow);
        c->d_grp = dupload(p.grp_v, ow);
        c->ngrp = (uint32_t)p.grp_v.size();
        c->d_n2e_dsp = dupload(p.n2e_dsp, ow);
        c->d_n2e_le = dupload(p.n2e_le, ow);
        c->d_n2e_pos = dupload(p.n2e_pos, ow);
        c->d_n2n_dsp = dupload(p.n2n_dsp, ow);
        c->d_n2n_lv = dupload(p.n2n_lv, ow);
        c->d_phi64 = dalloc<double>(c->ncopy, ow);
        c->d_st = dalloc<int8_t>(c->ncopy, ow);
        c->d_pending = dalloc<uint32_t>(1, ow);
        c->d_nexte = dalloc<double>(c->nsub, ow);
        c->d_tmpd = dalloc<double>(c->nn, ow);
        c->d_tmpi = dalloc<int32_t>(c->nn, ow);
        c->lds_conv = 0;
        c->xa = CompatArgs{c->d_sub, c->d_e2n, c->d_geo64, c->d_n2e_dsp, c->d_n2e_le, c->d_n2e_pos,
                           c->d_n2n_dsp, c->d_n2n_lv, c->d_phi64, c->d_st, c->d_stats,
                           c->d_pending, c->d_nexte, c->nsub, o->local == FIM_LOCAL_EXACT ? 1 : 0,
                           0, 0, 0, 0, o->apd > 0 ? o->apd : 3e5,
                           o_tphi, o_st, o_tst, o_snap, o_misc};
      }
      set_lds_attr(c);
      HIPCHECK(hipStreamSynchronize(c->stream));
      *out = c;
    } catch (...) {
      destroy_ctx(c);
      throw;
    }
    if (fim_reset(*out) != FIM_OK) {
      const std::string msg = fim_last_error();
      destroy_ctx(*out);
      *out = nullptr;
      throw Error(FIM_EDEVICE, msg);
    }
  });
}

int fim_reset(fim_ctx *c) {
  return guarded([&] {
    if (!c) throw Error(FIM_EINVAL, "null context");
    HIPCHECK(hipSetDevice(c->dev));
    if (c->o.mode == FIM_MODE_CONVERGE) {
      const size_t ncs = (size_t)c->ninst * c->ncopy, nss = (size_t)c->ninst * c->nsub;
      k_fill<float><<<blocks_for(ncs, 256), 256, 0, c->stream>>>(c->d_phi, ncs, F_INF);
      k_fill<uint8_t><<<blocks_for(ncs, 256), 256, 0, c->stream>>>(c->d_dirty, ncs, (uint8_t)0);
      k_fill<uint32_t><<<blocks_for(nss, 256), 256, 0, c->stream>>>(c->d_act, nss, 0u);
      k_fill<float><<<blocks_for(nss, 256), 256, 0, c->stream>>>(c->d_wake, nss, F_INF);
      k_fill<uint32_t><<<1, 64, 0, c->stream>>>(c->d_cnt, 64, 0u);
      k_fill<uint32_t><<<blocks_for(c->ninst, 256), 256, 0, c->stream>>>(c->d_start, c->ninst, 0u);
      for (int k = 0; k < 3; ++k)
        k_fill<uint32_t><<<blocks_for(c->nsub, 256), 256, 0, c->stream>>>(c->wl[k].bcnt, c->nsub, 0u);
    } else {
      k_fill<double><<<blocks_for(c->ncopy, 256), 256, 0, c->stream>>>(c->d_phi64, c->ncopy, DBL_MAX);
      k_fill<int8_t><<<blocks_for(c->ncopy, 256), 256, 0, c->stream>>>(c->d_st, c->ncopy, (int8_t)FIM_FAR);
      k_fill<double><<<blocks_for(c->nsub, 256), 256, 0, c->stream>>>(c->d_nexte, c->nsub, -INFINITY);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

namespace {
// seeds (instance, copy) pairs at t0 through the context's staging buffers
// (grown, never shrunk); returns once the pinned buffer may be reused
void seed_pairs(fim_ctx *c, const std::vector<uint2> &ic, double t0) {
  if (ic.empty()) return;
  if (c->o.mode == FIM_MODE_CONVERGE && !(t0 >= 0.0)) throw Error(FIM_EINVAL, "converge mode needs t0 >= 0");
  if (ic.size() >= (1ull << 32)) throw Error(FIM_ELIMIT, "too many seed copies");
  HIPCHECK(hipSetDevice(c->dev));
  if (ic.size() > c->seed_cap) {
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (c->h_seed) HIPCHECK(hipHostFree(c->h_seed));
    if (c->d_seed) HIPCHECK(hipFree(c->d_seed));
    c->h_seed = nullptr;
    c->d_seed = nullptr;
    c->seed_cap = 0;
    const size_t cap = std::max<size_t>(ic.size(), 1024);
    HIPCHECK(hipHostMalloc((void **)&c->h_seed, cap * sizeof(uint2), hipHostMallocDefault));
    HIPCHECK(hipMalloc(&c->d_seed, cap * sizeof(uint2)));
    c->seed_cap = cap;
  }
  std::memcpy(c->h_seed, ic.data(), ic.size() * sizeof(uint2));
  HIPCHECK(hipMemcpyAsync(c->d_seed, c->h_seed, ic.size() * sizeof(uint2), hipMemcpyHostToDevice, c->stream));
  const int g = blocks_for(ic.size(), 256);
  if (c->o.mode == FIM_MODE_CONVERGE) {
    k_conv_seed<<<g, 256, 0, c->stream>>>(c->d_seed, (uint32_t)ic.size(), (float)t0, c->d_c2s, c->d_phi,
                                          c->d_dirty, c->d_act, c->d_wake, c->wl[0], c->nsub,
                                          c->ncopy, c->ninst);
  } else {
    k_compat_seed<<<g, 256, 0, c->stream>>>(c->d_seed, (uint32_t)ic.size(), t0, c->d_c2s, c->d_phi64,
                                            c->d_st, c->d_nexte);
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c->stream));
}

void append_copies(const fim_ctx *c, int32_t inst, const int64_t *vtx, int64_t n, std::vector<uint2> &out) {
  for (uint32_t cp : copies_of(c, vtx, n)) out.push_back(make_uint2((uint32_t)inst, cp));
}
}  // namespace

int fim_seed(fim_ctx *c, int32_t inst, const int64_t *vtx, int64_t n, double t0) {
  return guarded([&] {
    if (!c || (!vtx && n > 0) || n < 0) throw Error(FIM_EINVAL, "bad seed arguments");
    check_inst(c, inst);
    std::vector<uint2> ic;
    append_copies(c, inst, vtx, n, ic);
    seed_pairs(c, ic, t0);
  });
}

int fim_seed_batch(fim_ctx *c, int32_t n_inst, const int32_t *inst, const int64_t *off,
                   const int64_t *vtx, double t0) {
  return guarded([&] {
    if (!c || n_inst < 0 || (n_inst > 0 && (!inst || !off))) throw Error(FIM_EINVAL, "bad seed arguments");
    if (n_inst == 0) return;
    if (off[0] != 0) throw Error(FIM_EINVAL, "off[0] must be 0");
    if (off[n_inst] > 0 && !vtx) throw Error(FIM_EINVAL, "bad seed arguments");
    std::vector<uint2> ic;
    for (int32_t i = 0; i < n_inst; ++i) {
      check_inst(c, inst[i]);
      if (off[i + 1] < off[i]) throw Error(FIM_EINVAL, "offsets must not decrease");
      append_copies(c, inst[i], vtx + off[i], off[i + 1] - off[i], ic);
    }
    seed_pairs(c, ic, t0);
  });
}

int fim_set_scaling(fim_ctx *c, int32_t inst, const float scale[4]) {
  return guarded([&] {
    if (!c || !scale) throw Error(FIM_EINVAL, "null argument");
    if (c->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "scaling needs converge mode");
    check_inst(c, inst);
    float m[4];
    for (int k = 0; k < 4; ++k) {
      if (!(scale[k] > 0.0f)) throw Error(FIM_EINVAL, "scalings must be positive");
      m[k] = 1.0f / (scale[k] * scale[k]);  // both speeds scaled: M scales by 1/s^2
    }
    std::memcpy(&c->h_rscale[4 * (size_t)inst], m, sizeof(m));
    c->params_dirty = true;
  });
}

int fim_set_window(fim_ctx *c, int32_t inst, double window) {
  return guarded([&] {
    if (!c) throw Error(FIM_EINVAL, "null context");
    if (c->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "time windows need converge mode");
    check_inst(c, inst);
    if (!(window >= 0.0) || !(window < 1e30)) throw Error(FIM_EINVAL, "window must be >= 0 and finite");
    c->h_winst[inst] = (float)window;
    c->params_dirty = true;
  });
}

int fim_init_field(fim_ctx *c, int32_t inst, const float *phi0) {
  return guarded([&] {
    if (!c || !phi0) throw Error(FIM_EINVAL, "null argument");
    if (c->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "warm start needs converge mode");
    check_inst(c, inst);
    for (uint32_t v = 0; v < c->nn; ++v)
      if (!(phi0[v] >= 0.0f)) throw Error(FIM_EINVAL, "initial field must be >= 0 (inf allowed)");
    HIPCHECK(hipSetDevice(c->dev));
    HIPCHECK(hipMemcpyAsync(c->d_tmpf, phi0, c->nn * sizeof(float), hipMemcpyHostToDevice, c->stream));
    k_conv_init<<<blocks_for(c->ncopy, 256), 256, 0, c->stream>>>(c->d_tmpf, c->d_c2g, (uint32_t)inst,
                                                                   c->d_phi, c->d_dirty, c->ncopy, 1);
    k_conv_queue_all<<<blocks_for(c->nsub, 256), 256, 0, c->stream>>>((uint32_t)inst, c->nsub, c->ninst,
                                                                       c->d_act, c->d_wake, c->wl[0]);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int fim_step(fim_ctx *c, double phi_start, double cur_time, double dt, double front_width,
             int32_t *n_active) {
  return guarded([&] {
    if (!c) throw Error(FIM_EINVAL, "null context");
    if (c->o.mode != FIM_MODE_COMPAT) throw Error(FIM_ESTATE, "fim_step needs compat mode");
    HIPCHECK(hipSetDevice(c->dev));
    HIPCHECK(hipEventRecord(c->ev0, c->stream));
    if (c->ngrp)
      k_compat_sync<<<blocks_for(c->ngrp, 256), 256, 0, c->stream>>>(
          c->d_grp, c->ngrp, c->d_g2c_dsp, c->d_g2c, c->d_c2s, c->d_phi64, c->d_st, c->d_nexte);
    k_fill<uint32_t><<<1, 64, 0, c->stream>>>(c->d_pending, 1, 0u);
    CompatArgs a = c->xa;
    a.fw = front_width;
    a.cur = cur_time;
    a.phi_start = phi_start;
    a.dt = dt;
    launch_compat_compute(c, a);
    HIPCHECK(hipEventRecord(c->ev1, c->stream));
    HIPCHECK(hipMemcpyAsync(c->h_pin, c->d_pending, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.last_ms = ms;
    if (n_active) *n_active = (int32_t)c->h_pin[0];
  });
}

namespace {
// The loop of global iterations in phases -- begin, a chunk of graph_iters
// iterations with the read-back of the queued units, collect, end -- so that
// one host thread can also drive several contexts at once (fim_run_multi),
// whose streams then overlap on the device.  Device counters accumulate from
// the last begin with reset_stats.
struct RunState {
  int32_t max_iters = 0, done = 0;
  uint32_t left = 1;
  double kms = 0.0;
  int64_t klaunch = 0;
};

void run_begin(fim_ctx *c, int32_t max_iters, bool reset_stats, RunState &r) {
  HIPCHECK(hipSetDevice(c->dev));
  flush_params(c);
  // profile mode launches the same sequence of kernels without the graph
  // (events cannot be timed inside a graph) and times every compute launch
  if (c->o.profile && c->kev.empty()) {
    c->kev.assign(2 * c->graph_iters, nullptr);
    for (auto &e : c->kev) HIPCHECK(hipEventCreate(&e));
  }
  if (!c->o.profile && !c->direct && !c->gexec) build_graph(c);
  if (reset_stats) {
    HIPCHECK(hipMemsetAsync(c->d_stats, 0, 8 * sizeof(unsigned long long), c->stream));
    c->stats.compute_ms = 0.0;
    c->stats.compute_launches = 0;
  }
  HIPCHECK(hipEventRecord(c->ev0, c->stream));
  // units of the list the seeds (or an earlier run) filled
  launch_plan(c, 0, -1);
  r = RunState{};
  r.max_iters = max_iters;
}

// enqueues graph_iters iterations, the read-back of the queued units and the
// chunk's completion event
void run_chunk(fim_ctx *c, RunState &r) {
  HIPCHECK(hipSetDevice(c->dev));
  if (c->o.profile || c->direct) {
    for (int i = 0; i < c->graph_iters; ++i) enqueue_iteration(c, i, c->o.profile != 0);
  } else {
    HIPCHECK(hipGraphLaunch(c->gexec, c->stream));
  }
  r.done += c->graph_iters;
  HIPCHECK(hipMemcpyAsync(c->h_pin, c->d_cnt, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipEventRecord(c->evc, c->stream));
}

// after the chunk's event has completed: launch times (profile mode) and
// queued units; true while more iterations are to run
bool run_collect(fim_ctx *c, RunState &r) {
  if (c->o.profile) {
    for (int j = 0; j < c->graph_iters; ++j) {
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, c->kev[2 * j], c->kev[2 * j + 1]));
      r.kms += ms;
      ++r.klaunch;
    }
  }
  r.left = c->h_pin[0];
  return r.left != 0 && r.done < r.max_iters;
}

void run_end(fim_ctx *c, RunState &r) {
  HIPCHECK(hipSetDevice(c->dev));
  HIPCHECK(hipEventRecord(c->ev1, c->stream));
  unsigned long long st[8];
  HIPCHECK(hipMemcpyAsync(st, c->d_stats, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->stats.last_ms = ms;
  c->stats.activations = (int64_t)st[0];
  c->stats.local_iters = (int64_t)st[1];
  c->stats.elem_updates = (int64_t)st[2];
  c->stats.global_iters = (int64_t)st[3];
  c->stats.alg_bytes = (int64_t)st[4];
  c->stats.vertex_updates = (int64_t)st[5];
  c->stats.compute_ms += r.kms;
  c->stats.compute_launches += r.klaunch;
}

// Runs global iterations in graph-sized chunks until no work is queued or
// max_iters (rounded up to whole chunks) have run; returns the queued units
// left (0 = every instance converged).
uint32_t run_iterations(fim_ctx *c, int32_t max_iters, bool reset_stats) {
  RunState r;
  run_begin(c, max_iters, reset_stats, r);
  do {
    run_chunk(c, r);
    HIPCHECK(hipEventSynchronize(c->evc));
  } while (run_collect(c, r));
  run_end(c, r);
  return r.left;
}

std::vector<uint32_t> checked_instances(const fim_ctx *c, int32_t n, const int32_t *insts) {
  if (n < 0 || (n > 0 && !insts)) throw Error(FIM_EINVAL, "bad instance list");
  std::vector<uint32_t> v(n);
  for (int32_t i = 0; i < n; ++i) {
    check_inst(c, insts[i]);
    v[i] = (uint32_t)insts[i];
  }
  return v;
}
}  // namespace

int fim_run(fim_ctx *c, int32_t max_iters, int32_t *iters) {
  return guarded([&] {
    if (!c) throw Error(FIM_EINVAL, "null context");
    if (c->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "fim_run needs converge mode");
    if (max_iters <= 0) throw Error(FIM_EINVAL, "max_iters must be positive");
    const uint32_t left = run_iterations(c, max_iters, true);
    if (iters) *iters = (int32_t)c->stats.global_iters;
    if (left != 0) throw Error(FIM_ENOCONV, "iteration cap reached before convergence");
  });
}

int fim_advance(fim_ctx *c, int32_t iters, int32_t reset_stats, int32_t *left) {
  return guarded([&] {
    if (!c) throw Error(FIM_EINVAL, "null context");
    if (c->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "fim_advance needs converge mode");
    if (iters <= 0) throw Error(FIM_EINVAL, "iters must be positive");
    const uint32_t l = run_iterations(c, iters, reset_stats != 0);
    if (left) *left = (int32_t)l;
  });
}

int fim_run_multi(fim_ctx *const *cs, int32_t n, int32_t max_iters, int32_t *iters) {
  return guarded([&] {
    if (!cs || n <= 0) throw Error(FIM_EINVAL, "bad context list");
    if (max_iters <= 0) throw Error(FIM_EINVAL, "max_iters must be positive");
    for (int32_t i = 0; i < n; ++i) {
      if (!cs[i]) throw Error(FIM_EINVAL, "null context");
      if (cs[i]->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "fim_run_multi needs converge mode");
      for (int32_t j = 0; j < i; ++j)
        if (cs[j] == cs[i]) throw Error(FIM_EINVAL, "context listed twice");
    }
    std::vector<RunState> rs(n);
    std::vector<uint8_t> running(n, 1);
    for (int32_t i = 0; i < n; ++i) {
      run_begin(cs[i], max_iters, true, rs[i]);
      run_chunk(cs[i], rs[i]);
    }
    // the host polls the chunks' events and keeps every stream fed; the
    // streams' kernels share the device, so one context's launch tail is
    // filled by another's work
    for (int32_t active = n; active > 0;) {
      bool progress = false;
      for (int32_t i = 0; i < n; ++i) {
        if (!running[i]) continue;
        HIPCHECK(hipSetDevice(cs[i]->dev));
        const hipError_t q = hipEventQuery(cs[i]->evc);
        if (q == hipErrorNotReady) continue;
        HIPCHECK(q);
        progress = true;
        if (run_collect(cs[i], rs[i])) {
          run_chunk(cs[i], rs[i]);
        } else {
          run_end(cs[i], rs[i]);
          running[i] = 0;
          --active;
        }
      }
      if (!progress) std::this_thread::yield();
    }
    bool capped = false;
    for (int32_t i = 0; i < n; ++i) {
      if (iters) iters[i] = (int32_t)cs[i]->stats.global_iters;
      capped = capped || rs[i].left != 0;
    }
    if (capped) throw Error(FIM_ENOCONV, "iteration cap reached before convergence");
  });
}

int fim_active(fim_ctx *c, int32_t *active) {
  return guarded([&] {
    if (!c || !active) throw Error(FIM_EINVAL, "null argument");
    if (c->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "fim_active needs converge mode");
    HIPCHECK(hipSetDevice(c->dev));
    k_inst_active<<<c->ninst, 256, 0, c->stream>>>(c->d_act, c->nsub, c->d_flag);
    HIPCHECK(hipGetLastError());
    std::vector<uint32_t> h(c->ninst);
    HIPCHECK(hipMemcpyAsync(h.data(), c->d_flag, sizeof(uint32_t) * c->ninst, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < c->ninst; ++i) active[i] = (int32_t)h[i];
  });
}

int fim_recycle(fim_ctx *c, int32_t n, const int32_t *insts) {
  return guarded([&] {
    if (!c) throw Error(FIM_EINVAL, "null context");
    if (c->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "fim_recycle needs converge mode");
    const std::vector<uint32_t> v = checked_instances(c, n, insts);
    if (v.empty()) return;
    HIPCHECK(hipSetDevice(c->dev));
    // only instances with nothing queued may be recycled
    k_inst_active<<<c->ninst, 256, 0, c->stream>>>(c->d_act, c->nsub, c->d_flag);
    HIPCHECK(hipGetLastError());
    std::vector<uint32_t> h(c->ninst);
    HIPCHECK(hipMemcpyAsync(h.data(), c->d_flag, sizeof(uint32_t) * c->ninst, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    std::vector<uint8_t> seen(c->ninst, 0);
    for (uint32_t i : v) {
      if (h[i]) throw Error(FIM_ESTATE, "instance " + std::to_string(i) + " still has queued work");
      if (seen[i]++) throw Error(FIM_EINVAL, "instance listed twice");
    }
    HIPCHECK(hipMemcpyAsync(c->d_flag, v.data(), sizeof(uint32_t) * v.size(), hipMemcpyHostToDevice, c->stream));
    const size_t total = ((size_t)c->ncopy + c->nsub) * v.size();
    k_inst_reset<<<blocks_for(total, 256), 256, 0, c->stream>>>(c->d_flag, (uint32_t)v.size(), c->d_phi, c->d_dirty,
                                                               c->d_wake, c->d_start, c->d_cnt + 48, c->ncopy,
                                                               c->nsub);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int fim_download(fim_ctx *c, int32_t inst, double *phi, int32_t *state) {
  return guarded([&] {
    if (!c) throw Error(FIM_EINVAL, "null context");
    check_inst(c, inst);
    HIPCHECK(hipSetDevice(c->dev));
    const int g = blocks_for(c->nn, 256);
    if (c->o.mode == FIM_MODE_CONVERGE) {
      k_conv_gather<<<g, 256, 0, c->stream>>>(c->d_phi, c->d_g2c_dsp, c->d_g2c, c->nn,
                                              (size_t)inst * c->ncopy, c->d_tmpf);
      HIPCHECK(hipGetLastError());
      std::vector<float> h(c->nn);
      HIPCHECK(hipMemcpyAsync(h.data(), c->d_tmpf, c->nn * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      HIPCHECK(hipStreamSynchronize(c->stream));
      for (uint32_t v = 0; v < c->nn; ++v) {
        const bool reached = h[v] < F_INF;
        if (phi) phi[v] = reached ? (double)h[v] : DBL_MAX;
        if (state) state[v] = reached ? FIM_FROZEN : FIM_FAR;
      }
    } else {
      k_compat_gather<<<g, 256, 0, c->stream>>>(c->d_phi64, c->d_st, c->d_g2c_dsp, c->d_g2c, c->nn,
                                                c->d_tmpd, c->d_tmpi);
      HIPCHECK(hipGetLastError());
      std::vector<double> hp(c->nn);
      std::vector<int32_t> hs(c->nn);
      HIPCHECK(hipMemcpyAsync(hp.data(), c->d_tmpd, c->nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHECK(hipMemcpyAsync(hs.data(), c->d_tmpi, c->nn * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
      HIPCHECK(hipStreamSynchronize(c->stream));
      if (phi) std::memcpy(phi, hp.data(), c->nn * sizeof(double));
      if (state) std::memcpy(state, hs.data(), c->nn * sizeof(int32_t));
    }
  });
}

int fim_residual(fim_ctx *c, int32_t inst, double *res) {
  return guarded([&] {
    if (!c || !res) throw Error(FIM_EINVAL, "null argument");
    if (c->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "residual needs converge mode");
    check_inst(c, inst);
    HIPCHECK(hipSetDevice(c->dev));
    flush_params(c);
    k_fill<uint32_t><<<1, 64, 0, c->stream>>>(c->d_tmpu, 1, 0u);
    k_conv_residual<<<(int)std::min<uint32_t>(c->nsub, 4096u), 256, 0, c->stream>>>(
        c->d_sub, c->nsub, c->d_e2n, c->d_geo32, c->d_ereg, c->d_rscale, c->d_phi,
        (size_t)inst * c->ncopy, (uint32_t)inst, c->d_tmpu);
    HIPCHECK(hipGetLastError());
    uint32_t bits = 0;
    HIPCHECK(hipMemcpyAsync(&bits, c->d_tmpu, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    float r;
    std::memcpy(&r, &bits, 4);
    *res = r;
  });
}

int fim_support(fim_ctx *c, int32_t inst, double *worst) {
  return guarded([&] {
    if (!c || !worst) throw Error(FIM_EINVAL, "null argument");
    if (c->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "support check needs converge mode");
    check_inst(c, inst);
    HIPCHECK(hipSetDevice(c->dev));
    uint32_t *best = nullptr;
    HIPCHECK(hipMalloc(&best, (size_t)c->nn * sizeof(uint32_t)));
    try {
      *worst = support_into(c, inst, best);
    } catch (...) {
      (void)hipFree(best);
      throw;
    }
    HIPCHECK(hipFree(best));
  });
}

int fim_support_noalloc(fim_ctx *c, int32_t inst, double *worst) {
  return guarded([&] {
    if (!c || !worst) throw Error(FIM_EINVAL, "null argument");
    if (c->o.mode != FIM_MODE_CONVERGE) throw Error(FIM_ESTATE, "support check needs converge mode");
    check_inst(c, inst);
    if (c->nn > kNoAllocMaxVertices)
      throw Error(FIM_ELIMIT, "fim_support_noalloc: " + std::to_string(c->nn) +
                                  " vertices exceed its scratch of " + std::to_string(kNoAllocMaxVertices) +
                                  " (use fim_support)");
    HIPCHECK(hipSetDevice(c->dev));
    *worst = support_into(c, inst, noalloc_scratch(c->dev));
  });
}

int fim_stats(const fim_ctx *c, fim_stats_t *s) {
  return guarded([&] {
    if (!c || !s) throw Error(FIM_EINVAL, "null argument");
    *s = c->stats;
  });
}

void fim_destroy(fim_ctx *c) {
  if (c) (void)hipSetDevice(c->dev);
  destroy_ctx(c);
}

int fim_device_info(int32_t device, char *buf, int64_t len) {
  return guarded([&] {
    if (!buf || len <= 0) throw Error(FIM_EINVAL, "null buffer");
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Error(FIM_EINVAL, "no such device");
    hipDeviceProp_t p;
    HIPCHECK(hipGetDeviceProperties(&p, device));
    std::snprintf(buf, (size_t)len,
                  "device %d: %s (%s), %d CUs, %.1f GiB, LDS/block %zu B, wavefront %d, clock %d kHz, "
                  "devices visible %d",
                  device, p.name, p.gcnArchName, p.multiProcessorCount,
                  (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0), (size_t)p.sharedMemPerBlock,
                  p.warpSize, p.clockRate, n);
  });
}

}  // extern "C"
