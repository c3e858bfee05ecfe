// oqh_internal.h -- shared state and device helpers of libopenqcd_hip.
//
// Device field layout ("site blocks"): every field keeps the reference's site
// order (ipt of lattice/geometry.c: even points first, then odd points, each
// half ordered by thread blocks, 2x2x2x2 sub-blocks and cache blocks).  The
// sites are grouped in blocks of 64 (one wavefront); inside a block each
// complex component is stored contiguously for the 64 sites:
//
//     unit(site, c) = ((site / 64) * NC + c) * 64 + site % 64
//
// with NC complex units per site: spinor 12, Pauli pair 36 (the 72 reals of
// swd[2x], swd[2x+1] taken as pairs), gauge 72 per odd site (8 links x 9,
// link order of udfld(): U(x,0), U(x-0,0), U(x,1), ...).  A wave therefore
// reads each component of 64 consecutive sites as one contiguous segment.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/openqcd_hip.h"

namespace oqh {

constexpr int SB = 64;  // sites per storage block (one wavefront)

__host__ __device__ inline size_t unit_index(int site, int c, int nc)
{
   return ((size_t)(site >> 6) * nc + c) * SB + (site & (SB - 1));
}

// number of complex units allocated for n sites with nc units per site
inline size_t units_for(int n, int nc) { return (size_t)((n + SB - 1) / SB) * nc * SB; }

enum Prec { F32 = 0, F64 = 1 };

struct Field
{
   void *base;
   int prec;
   int n;        // sites
   size_t bytes; // allocation size
};

struct Geometry
{
   int L[4] = {0, 0, 0, 0}, Ltrd[4] = {0, 0, 0, 0};
   int V = 0, VT = 0, nthread = 0;
   std::vector<int> ipt, lex, x0, iup, idn;  // iup/idn: [V][4]
};

struct Lib
{
   bool inited = false;
   int device = 0;
   hipStream_t stream = nullptr;
   bool abort_on_error = true;
   int nerr = 0;
   std::string errmsg;

   Geometry g;
   int bc = 3;
   double cF[2] = {1.0, 1.0};
   double m0 = 0.0, csw = 0.0;
   int eoflg = 0;
   int sap_bs[4] = {0, 0, 0, 0}, sap_isolv = 0, sap_nmr = 0, sap_ncy = 0;
   int sap_prec = 0;  // SAP storage: 0 FP32, 1 FP16 links, SW and spinors (per-site scales)
   int sap_graph = 1; // FP16 preconditioner of sap_gcr replayed as a HIP graph
   int gcr_cgs = 1;
   int nn_prec = 0;   // network correction in the preconditioner of sap_gcr   // fgcr orthogonalisation on the device: 1 two-pass (compact WY) MGS, 0 per-term MGS

   // geometry tables on the device
   int *d_nbr = nullptr;            // [8][V]: x+mu at 2mu, x-mu at 2mu+1; bit 31: boundary point
   unsigned char *d_bnd = nullptr;  // [V]: 1 on boundary time slices (open/SF bc rules)
   int *d_x0 = nullptr;             // [V]: time coordinate

   // gauge and SW fields
   double2 *d_ud = nullptr;   // 72 units per odd site
   float2 *d_u = nullptr;
   double2 *d_swd = nullptr;  // 36 units per site
   float2 *d_sw = nullptr;
   bool ud_set = false, u_valid = false, swd_valid = false, sw_valid = false;
   bool ud_phase = false;  // quark phase of set_ud_phase() applied to d_ud
   bool swd_inv[2] = {false, false};
   double2 *d_sws = nullptr;  // SW term as S1, S2, S3 per half (27 units per point)
   bool sws_valid = false;    // d_sws holds the term that compute_swd last wrote
   float2 *d_sws32 = nullptr; // its FP32 rounding
   bool sws32_valid = false;
   int swd_epoch = 0;  // incremented whenever the contents of d_swd change
   int u_epoch = 0;    // incremented whenever d_u is remade
   int ud_version = 0; // incremented whenever d_ud changes
   int rec_version = -1;            // ud_version the link classes were made for
   bool ud_rec = false;             // every link equals +-(rows 0,1, conj(row0 x row1))
   double ud_rec_dev = -1.0;        // largest deviation measured for that test
   unsigned char *d_lsign = nullptr; // per odd point: bit l set if link l is -SU(3)

   std::map<uintptr_t, Field> reg;

   // staging for host-array arguments
   void *stage[4] = {nullptr, nullptr, nullptr, nullptr};
   size_t stage_bytes[4] = {0, 0, 0, 0};
   void *d_tmp = nullptr;  // device scratch (odd-site intermediate of Dwhat, reductions)
   size_t tmp_bytes = 0;

   // workspaces (reference wspace.c semantics)
   std::vector<spinor *> ws;
   std::vector<spinor_dble *> wsd;
   int ws_used = 0, wsd_used = 0;
   std::vector<spinor *> ws_ptrs;
   std::vector<spinor_dble *> wsd_ptrs;
};

Lib &lib();

// error handling (abort like error() in the reference, or record)
void fail(const char *where, const char *msg);
#define OQH_CHECK(cond, where, msg) \
   do { if (!(cond)) { oqh::fail(where, msg); } } while (0)
#define HIP_OK(call)                                                         \
   do {                                                                      \
      hipError_t e_ = (call);                                                \
      if (e_ != hipSuccess) oqh::fail(#call, hipGetErrorString(e_));         \
   } while (0)

bool require_geometry(const char *where);
void *dev_alloc(size_t bytes);
void dev_free(void *p);
void *tmp_buffer(size_t bytes);

// field resolution: device base pointer (and site offset) of an argument
struct Res
{
   void *dev = nullptr;  // device base of the field (site 0)
   int off = 0;          // site offset addressed by the pointer
   int n = 0;            // sites available from off
   bool host = false;    // argument is a host array (staged)
   void *hostptr = nullptr;
   int slot = -1;
};
Res resolve(const void *p, int prec, int nsites, int slot, bool upload);
void *device_field(const void *p, int prec, int *off);
void *device_field_n(const void *p, int prec, int *off, int *avail);
void finish(const Res &r, int prec, int nsites, bool download);

// geometry and tables
void build_geometry(Geometry &g, const int *L, const int *Ltrd);
void upload_tables();
void sap_release();  // frees the SAP block grid (made for the current geometry)
void release_flow(); // frees the force field of the Wilson flow
void release_nn();   // frees the network of the learned preconditioner
void release_rng();  // frees the device random number streams
void release_dfl();  // frees the deflation subspace and resets its parameters
bool nn_ready();     // network created and enabled in sap_gcr
bool nn_correct(const float2 *r, float2 *phi);  // phi += NN(r)

// layout conversions (reference AoS <-> site blocks), n sites from site 0
void aos_to_blocks_f64(const double2 *aos, double2 *blk, int n, int nc, hipStream_t st);
void blocks_to_aos_f64(const double2 *blk, double2 *aos, int n, int nc, hipStream_t st);
void aos_to_blocks_f32(const float2 *aos, float2 *blk, int n, int nc, hipStream_t st);
void blocks_to_aos_f32(const float2 *blk, float2 *aos, int n, int nc, hipStream_t st);
void convert_f64_to_f32(const double2 *a, float2 *b, size_t nunits, hipStream_t st);

// fused Gram-Schmidt steps of fgcr on single-precision device fields of n
// points (linalg.hip); coefficients in device slots coef[ic] (input) and
// coef[io] (output)
enum GsMode { GS_DOT = 0, GS_AXPY_DOT, GS_AXPY_NORM, GS_SCALE_DOT, GS_NORM };
template <int MODE>
void gs_step(float2 *s, const float2 *r, const float2 *q, float2 *coef, int ic, int io, int n);
// Gram-Schmidt step of fgcr in two passes (compact WY form of modified GS):
// coef[j] = d_j (the modified-GS coefficients) for j < k, s -= sum_j d_j q_j,
// coef[ib] = (|s|, |s|^2), s /= |s|, coef[ic] = (s, rho), rho -= coef[ic] s,
// coef[ir] = (|rho|, |rho|^2); the Gram entries of q[k-1] are formed on the
// way into gram[(k-1) ldg + j] (q[0..k-2] must be the vectors of the
// previous calls with the same gram; ldg >= k)
void gs_classical(float2 *s, float2 *const *q, int k, float2 *rho, float2 *coef, int ib, int ic, int ir, int n,
                  float2 *gram, int ldg);

// SW field management (swterm.hip)
void ensure_swd(int want_inv_e, int want_inv_o);
int ensure_swd_ret(int want_inv_e, int want_inv_o);
int invert_pauli(const double2 *in, double2 *out, int p0, int np, double mu);
void ensure_sw();
void ensure_u();

}  // namespace oqh

// ---------------------------------------------------------------------------
// complex arithmetic on double2 / float2
// ---------------------------------------------------------------------------
template <typename T> struct cplx_of;
template <> struct cplx_of<double> { typedef double2 t; };
template <> struct cplx_of<float> { typedef float2 t; };

template <typename C> __host__ __device__ inline C cmk(decltype(C::x) a, decltype(C::x) b)
{
   C c;
   c.x = a;
   c.y = b;
   return c;
}
template <typename C> __device__ inline C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ inline C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
// a + i b, a - i b
template <typename C> __device__ inline C caddi(C a, C b) { return cmk<C>(a.x - b.y, a.y + b.x); }
template <typename C> __device__ inline C csubi(C a, C b) { return cmk<C>(a.x + b.y, a.y - b.x); }
template <typename C> __device__ inline C cmul(C a, C b)
{
   return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
template <typename C> __device__ inline C cmulc(C a, C b)
{
   return cmk<C>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename C> __device__ inline C cscale(C a, decltype(C::x) s) { return cmk<C>(a.x * s, a.y * s); }
