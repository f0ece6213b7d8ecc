// host.hpp -- host-side data structures of the FIM solver (mesh, partition,
// device pack).  Internal to libcufim.so; the public surface is include/cufim.h.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "cufim.h"

namespace fim {

// error carried to the C ABI as (code, message)
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

// ---------------------------------------------------------------------------
// Mesh: tetrahedra only.  Element node lists are kept sorted ascending, the
// order the reference establishes in compute_base_connectivity
// (src/core/types.cpp:63-67) and writes to its VTK output.
// ---------------------------------------------------------------------------
struct Mesh {
  int64_t nn = 0, ne = 0;
  std::vector<double> xyz;      // 3*nn
  std::vector<int64_t> e2n;     // 4*ne, sorted per element
  std::vector<double> fib;      // 3*ne unit fibre
  std::vector<double> vel;      // 2*ne (v_fibre, v_cross)
  std::vector<int32_t> etags;   // ne, written as elemTag
  std::vector<int32_t> atags;   // ne partition tags (may be empty)
  std::vector<uint8_t> region;  // ne region ids 0..3 (may be empty)

  // vertex -> incident elements, ascending element id
  std::vector<int64_t> n2e_dsp;  // nn+1
  std::vector<int64_t> n2e_con;  // 4*ne

  void finalize();  // sorts element nodes, validates, builds n2e
};

// metric tensor M = f f^T / vf^2 + (I - f f^T) / vs^2 packed (xx xy xz yy yz zz)
void metric_tensor(const double f[3], double vf, double vs, double M[6]);
// 6 squared metric edge lengths e12 e13 e23 e14 e24 e34 of element e in the
// element's (sorted) node order
void element_geo(const Mesh &m, int64_t e, double geo[6]);
double front_width(const Mesh &m);

Mesh load_vtk(const std::string &path);
void save_vtk(const Mesh &m, const std::string &path, const double *phi,
              const int32_t *state);
std::vector<int32_t> read_tags(const std::string &path);
Mesh gen_box(double lx, double ly, double lz, int nx, int ny, int nz, double jitter,
             uint64_t seed, double a0, double a1, double vf, double vc);
Mesh gen_shell(double ao, double bo, double ai, double bi, double h, int n_scars, double r0,
               double r1, uint64_t seed, double vf, double vc, double vs);

// ---------------------------------------------------------------------------
// Partition: elements grouped into submeshes, vertices duplicated per
// submesh ("copies"), copies of one submesh contiguous.
// ---------------------------------------------------------------------------
struct Partition {
  int64_t nsub = 0;
  std::vector<int64_t> e2s;        // ne
  std::vector<int64_t> s_edsp;     // nsub+1 offsets into s_elems
  std::vector<int64_t> s_elems;    // element ids in submesh order
  std::vector<int64_t> s_vdsp;     // nsub+1 offsets into l2g
  std::vector<int64_t> l2g;        // copy -> global vertex
};

Partition partition_from_tags(const Mesh &m, const std::vector<int32_t> &tags);
Partition partition_rcb(const Mesh &m, int max_nn, int max_ne, uint64_t seed);

// ---------------------------------------------------------------------------
// Pack: device layout.  All offsets are absolute indices into the flat
// arrays below.
// ---------------------------------------------------------------------------
struct SubInfo {
  uint32_t eoff, ne;   // elements [eoff, eoff+ne)
  uint32_t coff, nn;   // copies   [coff, coff+nn)
  uint32_t ioff, ni;   // interface entries [ioff, ioff+ni)
};

struct Pack {
  int64_t nn_global = 0, ne_global = 0, nsub = 0, ncopy = 0;
  int32_t nn_max = 0, ne_max = 0;
  int32_t source = FIM_PART_RCB;
  Partition part;
  std::vector<SubInfo> sub;        // nsub
  std::vector<uint16_t> e2n_l;     // 4 per element: local copy ids, sorted
  std::vector<float> geo32;        // 6 per element
  std::vector<float> hgt32;        // 4 per element: metric height of each node
                                   // above its opposite face, rounded down
  std::vector<double> geo64;       // 6 per element
  std::vector<uint8_t> ereg;       // 1 per element
  std::vector<uint32_t> c2g;       // per copy
  std::vector<uint32_t> c2s;       // per copy
  std::vector<uint32_t> g2c_dsp;   // nn_global+1
  std::vector<uint32_t> g2c;       // ncopy, ascending submesh
  std::vector<uint32_t> grp_v;     // global vertices with >= 2 copies
  // interface (converge mode): per submesh entries [ioff, ioff+ni)
  std::vector<uint16_t> if_l;      // local copy id of the entry
  std::vector<uint32_t> if_dsp;    // n_if+1 offsets into if_other
  std::vector<uint32_t> if_other;  // other copies (global copy ids)
  // compat extras: per copy CSR into local element / vertex lists
  std::vector<uint32_t> n2e_dsp;   // ncopy+1
  std::vector<uint16_t> n2e_le;    // local element id
  std::vector<uint8_t> n2e_pos;    // position of the owner in that element
  std::vector<uint32_t> n2n_dsp;   // ncopy+1
  std::vector<uint16_t> n2n_lv;    // local copy ids, sorted, self included
  double front_width = 0.0;
};

Pack build_pack(const Mesh &m, const Partition &p, int32_t source);

// element-list entries the converge-mode band kernel keeps per pass (more
// marked elements take further passes)
constexpr int32_t kElistCap = 2048;

// dynamic LDS bytes the converge-mode kernel needs for the largest submesh
int64_t converge_lds_bytes(int32_t nn_max, int32_t ne_max);

// solver-context hooks for the surrogate (csrc/surrogate.hip); device
// pointers, work queued on the context's stream
void *ctx_stream(fim_ctx *c);
int32_t ctx_device(const fim_ctx *c);
uint32_t ctx_vertices(const fim_ctx *c);
int32_t ctx_instances(const fim_ctx *c);
bool ctx_converge(const fim_ctx *c);
// per-vertex map of an instance (min over copies) in the context's field buffer
const float *ctx_gather(fim_ctx *c, int32_t inst);
// the context's field buffer (nn floats): fill it, then ctx_init_field
float *ctx_field_buffer(fim_ctx *c);
// instance field := field buffer; all_active marks every copy dirty and
// queues every submesh (warm start), otherwise only the times are set
void ctx_init_field(fim_ctx *c, int32_t inst, bool all_active);

}  // namespace fim

struct fim_mesh {
  fim::Mesh m;
};
struct fim_pack {
  fim::Pack p;
};
