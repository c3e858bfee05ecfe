// state.hip -- library state, lattice geometry, field registry, layout
// conversion, import/export and the spinor workspaces.
//
// Geometry: the site order ipt is the one of lattice/geometry.c:50-190
// (thread blocks of size L*_TRD, each split into 2x2x2x2 sub-blocks of sizes
// L_TRD/2+(L_TRD/2)%2 and the rest, each split into cache blocks whose time
// extent is the sub-block extent and whose space extents are 4, 3 or 2;
// lexicographic order inside, even points numbered from 0 and odd points
// from VOLUME/2).  Keeping this order makes every reference pointer offset
// valid on the device fields.  Workspaces follow utils/wspace.c:394-510.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>

#include "oqh_internal.h"

namespace oqh {

Lib &lib()
{
   static Lib L;
   return L;
}

void fail(const char *where, const char *msg)
{
   Lib &L = lib();
   char buf[512];
   snprintf(buf, sizeof(buf), "Error in %s:\n%s", where, msg);
   L.nerr += 1;
   L.errmsg = buf;
   if (L.abort_on_error)
   {
      fprintf(stderr, "\n%s\nProgram aborted\n\n", buf);
      fflush(stderr);
      abort();
   }
}

bool require_geometry(const char *where)
{
   Lib &L = lib();
   if (!L.inited) oqh_init(0);
   if (L.g.V == 0)
   {
      fail(where, "lattice geometry is not set (oqh_geometry)");
      return false;
   }
   return true;
}

void *dev_alloc(size_t bytes)
{
   void *p = nullptr;
   if (bytes == 0) bytes = 256;
   hipError_t e = hipMalloc(&p, bytes);
   if (e != hipSuccess)
   {
      fail("dev_alloc", hipGetErrorString(e));
      return nullptr;
   }
   return p;
}

void dev_free(void *p)
{
   if (p) (void)hipFree(p);
}

void *tmp_buffer(size_t bytes)
{
   Lib &L = lib();
   if (bytes > L.tmp_bytes)
   {
      if (L.d_tmp)
      {
         HIP_OK(hipStreamSynchronize(L.stream));
         dev_free(L.d_tmp);
      }
      L.d_tmp = dev_alloc(bytes);
      L.tmp_bytes = bytes;
   }
   return L.d_tmp;
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------

static int lexi(const int *L, const int *x) { return x[3] + L[3] * (x[2] + L[2] * (x[1] + L[1] * x[0])); }

static void number_box(Geometry &g, const int *org, const int *ext, int *ofs)
{
   int cb[4], nb[4], j[4], x[4], y[4];
   cb[0] = ext[0];
   for (int mu = 1; mu < 4; mu++) cb[mu] = (ext[mu] % 4 == 0) ? 4 : ((ext[mu] % 3 == 0) ? 3 : 2);
   for (int mu = 0; mu < 4; mu++) nb[mu] = ext[mu] / cb[mu];
   for (j[0] = 0; j[0] < nb[0]; j[0]++)
   for (j[1] = 0; j[1] < nb[1]; j[1]++)
   for (j[2] = 0; j[2] < nb[2]; j[2]++)
   for (j[3] = 0; j[3] < nb[3]; j[3]++)
      for (x[0] = 0; x[0] < cb[0]; x[0]++)
      for (x[1] = 0; x[1] < cb[1]; x[1]++)
      for (x[2] = 0; x[2] < cb[2]; x[2]++)
      for (x[3] = 0; x[3] < cb[3]; x[3]++)
      {
         for (int mu = 0; mu < 4; mu++) y[mu] = org[mu] + j[mu] * cb[mu] + x[mu];
         g.ipt[lexi(g.L, y)] = ofs[(y[0] + y[1] + y[2] + y[3]) & 1]++;
      }
}

void build_geometry(Geometry &g, const int *L, const int *Ltrd)
{
   int sb[2][4], nt[4];
   for (int mu = 0; mu < 4; mu++)
   {
      g.L[mu] = L[mu];
      g.Ltrd[mu] = Ltrd[mu];
      sb[0][mu] = Ltrd[mu] / 2 + (Ltrd[mu] / 2) % 2;
      sb[1][mu] = Ltrd[mu] - sb[0][mu];
      nt[mu] = L[mu] / Ltrd[mu];
   }
   g.V = L[0] * L[1] * L[2] * L[3];
   g.VT = Ltrd[0] * Ltrd[1] * Ltrd[2] * Ltrd[3];
   g.nthread = g.V / g.VT;
   g.ipt.assign(g.V, 0);
   g.lex.assign(g.V, 0);
   g.x0.assign(g.V, 0);
   g.iup.assign(4 * (size_t)g.V, 0);
   g.idn.assign(4 * (size_t)g.V, 0);

   int ofs[2] = {0, g.V / 2}, i[4], org[4], ext[4];
   for (i[0] = 0; i[0] < 2; i[0]++)
   for (i[1] = 0; i[1] < 2; i[1]++)
   for (i[2] = 0; i[2] < 2; i[2]++)
   for (i[3] = 0; i[3] < 2; i[3]++)
   {
      for (int mu = 0; mu < 4; mu++)
      {
         ext[mu] = sb[i[mu]][mu];
         org[mu] = i[mu] * sb[0][mu];
      }
      number_box(g, org, ext, ofs);
   }
   for (int k = 1; k < g.nthread; k++)
   {
      int n = k, b[4], x[4], y[4];
      b[3] = n % nt[3]; n /= nt[3];
      b[2] = n % nt[2]; n /= nt[2];
      b[1] = n % nt[1]; n /= nt[1];
      b[0] = n;
      for (x[0] = 0; x[0] < Ltrd[0]; x[0]++)
      for (x[1] = 0; x[1] < Ltrd[1]; x[1]++)
      for (x[2] = 0; x[2] < Ltrd[2]; x[2]++)
      for (x[3] = 0; x[3] < Ltrd[3]; x[3]++)
      {
         for (int mu = 0; mu < 4; mu++) y[mu] = x[mu] + b[mu] * Ltrd[mu];
         g.ipt[lexi(L, y)] = g.ipt[lexi(L, x)] + k * (g.VT / 2);
      }
   }
   for (int lx = 0; lx < g.V; lx++) g.lex[g.ipt[lx]] = lx;
   for (int ix = 0; ix < g.V; ix++)
   {
      int lx = g.lex[ix], x[4], y[4];
      x[3] = lx % L[3];
      x[2] = (lx / L[3]) % L[2];
      x[1] = (lx / (L[3] * L[2])) % L[1];
      x[0] = lx / (L[3] * L[2] * L[1]);
      g.x0[ix] = x[0];
      for (int mu = 0; mu < 4; mu++)
      {
         memcpy(y, x, sizeof(y));
         y[mu] = (x[mu] + 1) % L[mu];
         g.iup[4 * (size_t)ix + mu] = g.ipt[lexi(L, y)];
         y[mu] = (x[mu] + L[mu] - 1) % L[mu];
         g.idn[4 * (size_t)ix + mu] = g.ipt[lexi(L, y)];
      }
   }
}

// boundary time slices where quark fields vanish (lattice/bcnds.c:275-354)
static inline bool bnd_point(const Lib &L, int ix)
{
   int t = L.g.x0[ix];
   return (t == 0 && L.bc != 3) || (t == L.g.L[0] - 1 && L.bc == 0);
}

void upload_tables()
{
   Lib &L = lib();
   int V = L.g.V;
   if (V == 0) return;
   std::vector<int> nbr(8 * (size_t)V);
   std::vector<unsigned char> bnd(V);
   for (int ix = 0; ix < V; ix++)
   {
      bnd[ix] = bnd_point(L, ix) ? 1 : 0;
      for (int mu = 0; mu < 4; mu++)
      {
         int yu = L.g.iup[4 * (size_t)ix + mu], yd = L.g.idn[4 * (size_t)ix + mu];
         nbr[(size_t)(2 * mu) * V + ix] = yu | (bnd_point(L, yu) ? (int)0x80000000u : 0);
         nbr[(size_t)(2 * mu + 1) * V + ix] = yd | (bnd_point(L, yd) ? (int)0x80000000u : 0);
      }
   }
   if (!L.d_nbr) L.d_nbr = (int *)dev_alloc(sizeof(int) * 8 * (size_t)V);
   if (!L.d_bnd) L.d_bnd = (unsigned char *)dev_alloc(V);
   if (!L.d_x0) L.d_x0 = (int *)dev_alloc(sizeof(int) * (size_t)V);
   HIP_OK(hipMemcpy(L.d_nbr, nbr.data(), sizeof(int) * 8 * (size_t)V, hipMemcpyHostToDevice));
   HIP_OK(hipMemcpy(L.d_bnd, bnd.data(), V, hipMemcpyHostToDevice));
   HIP_OK(hipMemcpy(L.d_x0, L.g.x0.data(), sizeof(int) * (size_t)V, hipMemcpyHostToDevice));
}

// ---------------------------------------------------------------------------
// layout conversion kernels
// ---------------------------------------------------------------------------

template <typename C>
__global__ void k_aos_to_blocks(const C *__restrict__ aos, C *__restrict__ blk, int off, int n, int nc)
{
   size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= (size_t)n * nc) return;
   int site = (int)(i / nc), c = (int)(i % nc);
   blk[unit_index(off + site, c, nc)] = aos[i];
}

template <typename C>
__global__ void k_blocks_to_aos(const C *__restrict__ blk, C *__restrict__ aos, int off, int n, int nc)
{
   size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= (size_t)n * nc) return;
   int site = (int)(i / nc), c = (int)(i % nc);
   aos[i] = blk[unit_index(off + site, c, nc)];
}

__global__ void k_f64_to_f32(const double2 *__restrict__ a, float2 *__restrict__ b, size_t n)
{
   size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

static unsigned nblk(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

void aos_to_blocks_f64(const double2 *aos, double2 *blk, int n, int nc, hipStream_t st)
{
   if (n > 0) k_aos_to_blocks<double2><<<nblk((size_t)n * nc), 256, 0, st>>>(aos, blk, 0, n, nc);
}
void blocks_to_aos_f64(const double2 *blk, double2 *aos, int n, int nc, hipStream_t st)
{
   if (n > 0) k_blocks_to_aos<double2><<<nblk((size_t)n * nc), 256, 0, st>>>(blk, aos, 0, n, nc);
}
void aos_to_blocks_f32(const float2 *aos, float2 *blk, int n, int nc, hipStream_t st)
{
   if (n > 0) k_aos_to_blocks<float2><<<nblk((size_t)n * nc), 256, 0, st>>>(aos, blk, 0, n, nc);
}
void blocks_to_aos_f32(const float2 *blk, float2 *aos, int n, int nc, hipStream_t st)
{
   if (n > 0) k_blocks_to_aos<float2><<<nblk((size_t)n * nc), 256, 0, st>>>(blk, aos, 0, n, nc);
}
void convert_f64_to_f32(const double2 *a, float2 *b, size_t nunits, hipStream_t st)
{
   if (nunits) k_f64_to_f32<<<nblk(nunits), 256, 0, st>>>(a, b, nunits);
}

// ---------------------------------------------------------------------------
// field registry and argument resolution
// ---------------------------------------------------------------------------

static size_t elem_bytes(int prec) { return prec == F64 ? sizeof(spinor_dble) : sizeof(spinor); }

static void *register_field(int prec, int n)
{
   Lib &L = lib();
   size_t bytes = units_for(n, 12) * (prec == F64 ? sizeof(double2) : sizeof(float2));
   // pointer arithmetic of the caller must stay inside the allocation
   bytes = std::max(bytes, (size_t)n * elem_bytes(prec));
   void *p = dev_alloc(bytes);
   if (!p) return nullptr;
   HIP_OK(hipMemsetAsync(p, 0, bytes, L.stream));
   L.reg[(uintptr_t)p] = Field{p, prec, n, bytes};
   return p;
}

static const Field *lookup(const void *p, int *off)
{
   Lib &L = lib();
   auto it = L.reg.upper_bound((uintptr_t)p);
   if (it == L.reg.begin()) return nullptr;
   --it;
   const Field &f = it->second;
   uintptr_t d = (uintptr_t)p - it->first;
   if (d >= (uintptr_t)f.n * elem_bytes(f.prec)) return nullptr;
   if (d % elem_bytes(f.prec)) return nullptr;
   *off = (int)(d / elem_bytes(f.prec));
   return &f;
}

// device base of a registered field of the given precision (nullptr if p is
// not one), with the site offset addressed by p
void *device_field(const void *p, int prec, int *off)
{
   int o = 0;
   const Field *f = p ? lookup(p, &o) : nullptr;
   if (!f || f->prec != prec) return nullptr;
   *off = o;
   return f->base;
}

// as device_field, with the number of sites available from the offset
void *device_field_n(const void *p, int prec, int *off, int *avail)
{
   int o = 0;
   const Field *f = p ? lookup(p, &o) : nullptr;
   if (!f || f->prec != prec) return nullptr;
   *off = o;
   *avail = f->n - o;
   return f->base;
}

Res resolve(const void *p, int prec, int nsites, int slot, bool upload)
{
   Lib &L = lib();
   Res r;
   if (p == nullptr)
   {
      fail("resolve", "NULL field pointer");
      return r;
   }
   int off = 0;
   const Field *f = lookup(p, &off);
   if (f)
   {
      if (f->prec != prec)
      {
         fail("resolve", "field has the wrong precision");
         return r;
      }
      r.dev = f->base;
      r.off = off;
      r.n = f->n - off;
      if (r.n < nsites) fail("resolve", "field is too short for the requested volume");
      return r;
   }
   // host array in the reference layout: stage it on the device
   size_t blkbytes = units_for(nsites, 12) * (prec == F64 ? sizeof(double2) : sizeof(float2));
   size_t need = 2 * blkbytes;
   if (L.stage_bytes[slot] < need)
   {
      HIP_OK(hipStreamSynchronize(L.stream));
      dev_free(L.stage[slot]);
      L.stage[slot] = dev_alloc(need);
      L.stage_bytes[slot] = need;
   }
   char *aos = (char *)L.stage[slot];
   void *blk = aos + blkbytes;
   r.dev = blk;
   r.off = 0;
   r.n = nsites;
   r.host = true;
   r.hostptr = (void *)p;
   r.slot = slot;
   if (upload)
   {
      HIP_OK(hipMemcpyAsync(aos, p, (size_t)nsites * elem_bytes(prec), hipMemcpyHostToDevice, L.stream));
      if (prec == F64)
         aos_to_blocks_f64((const double2 *)aos, (double2 *)blk, nsites, 12, L.stream);
      else
         aos_to_blocks_f32((const float2 *)aos, (float2 *)blk, nsites, 12, L.stream);
   }
   else
      HIP_OK(hipMemsetAsync(blk, 0, blkbytes, L.stream));
   return r;
}

void finish(const Res &r, int prec, int nsites, bool download)
{
   Lib &L = lib();
   if (!r.host || !download) return;
   char *aos = (char *)L.stage[r.slot];
   if (prec == F64)
      blocks_to_aos_f64((const double2 *)r.dev, (double2 *)aos, nsites, 12, L.stream);
   else
      blocks_to_aos_f32((const float2 *)r.dev, (float2 *)aos, nsites, 12, L.stream);
   HIP_OK(hipMemcpyAsync(r.hostptr, aos, (size_t)nsites * elem_bytes(prec), hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}

static void import_field(void *dev, const void *host, int n, int prec)
{
   Lib &L = lib();
   int off = 0;
   const Field *f = lookup(dev, &off);
   if (!f || f->prec != prec || off + n > f->n)
   {
      fail("oqh_import", "not a device field of this library (or out of range)");
      return;
   }
   size_t eb = elem_bytes(prec);
   void *aos = tmp_buffer((size_t)n * eb);
   HIP_OK(hipMemcpyAsync(aos, host, (size_t)n * eb, hipMemcpyHostToDevice, L.stream));
   unsigned nb = nblk((size_t)n * 12);
   if (prec == F64)
      k_aos_to_blocks<double2><<<nb, 256, 0, L.stream>>>((const double2 *)aos, (double2 *)f->base, off, n, 12);
   else
      k_aos_to_blocks<float2><<<nb, 256, 0, L.stream>>>((const float2 *)aos, (float2 *)f->base, off, n, 12);
   HIP_OK(hipStreamSynchronize(L.stream));
}

static void export_field(const void *dev, void *host, int n, int prec)
{
   Lib &L = lib();
   int off = 0;
   const Field *f = lookup(dev, &off);
   if (!f || f->prec != prec || off + n > f->n)
   {
      fail("oqh_export", "not a device field of this library (or out of range)");
      return;
   }
   size_t eb = elem_bytes(prec);
   void *aos = tmp_buffer((size_t)n * eb);
   unsigned nb = nblk((size_t)n * 12);
   if (prec == F64)
      k_blocks_to_aos<double2><<<nb, 256, 0, L.stream>>>((const double2 *)f->base, (double2 *)aos, off, n, 12);
   else
      k_blocks_to_aos<float2><<<nb, 256, 0, L.stream>>>((const float2 *)f->base, (float2 *)aos, off, n, 12);
   HIP_OK(hipMemcpyAsync(host, aos, (size_t)n * eb, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}

}  // namespace oqh

using namespace oqh;

// ---------------------------------------------------------------------------
// C ABI: part A
// ---------------------------------------------------------------------------

extern "C" int oqh_init(int device)
{
   Lib &L = lib();
   if (L.inited) return 0;
   int n = 0;
   if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
   {
      fail("oqh_init", "no HIP device available");
      return 1;
   }
   if (device < 0 || device >= n) device = 0;
   HIP_OK(hipSetDevice(device));
   HIP_OK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
   L.device = device;
   L.inited = true;
   return 0;
}

static void clear_ws_marks();

extern "C" void oqh_finalize(void)
{
   Lib &L = lib();
   if (!L.inited) return;
   (void)hipStreamSynchronize(L.stream);
   sap_release();
   release_flow();
   clear_ws_marks();
   L.sap_prec = 0;
   L.gcr_cgs = 1;
   dev_free(L.d_lsign);
   L.d_lsign = nullptr;
   L.rec_version = -1;
   L.ud_rec = false;
   release_nn();
   release_rng();
   release_dfl();
   L.nn_prec = 0;
   L.sap_graph = 1;
   for (auto &kv : L.reg) dev_free(kv.second.base);
   L.reg.clear();
   for (int k = 0; k < 4; k++)
   {
      dev_free(L.stage[k]);
      L.stage[k] = nullptr;
      L.stage_bytes[k] = 0;
   }
   dev_free(L.d_tmp);
   L.d_tmp = nullptr;
   L.tmp_bytes = 0;
   dev_free(L.d_nbr); dev_free(L.d_bnd); dev_free(L.d_x0);
   dev_free(L.d_ud); dev_free(L.d_u); dev_free(L.d_swd); dev_free(L.d_sw); dev_free(L.d_sws); dev_free(L.d_sws32);
   L.d_sws = nullptr;
   L.d_sws32 = nullptr;
   L.sws_valid = L.sws32_valid = false;
   L.d_nbr = nullptr; L.d_bnd = nullptr; L.d_x0 = nullptr;
   L.d_ud = nullptr; L.d_u = nullptr; L.d_swd = nullptr; L.d_sw = nullptr;
   L.ws.clear(); L.wsd.clear(); L.ws_ptrs.clear(); L.wsd_ptrs.clear();
   L.ws_used = L.wsd_used = 0;
   L.ud_set = L.u_valid = L.swd_valid = L.sw_valid = false;
   (void)hipStreamDestroy(L.stream);
   L.stream = nullptr;
   L.g = Geometry();
   L.inited = false;
}

extern "C" void oqh_set_error_mode(int abort_on_error) { lib().abort_on_error = abort_on_error != 0; }

extern "C" int oqh_last_error(char *msg, int len)
{
   Lib &L = lib();
   int n = L.nerr;
   if (msg && len > 0)
   {
      strncpy(msg, L.errmsg.c_str(), (size_t)len - 1);
      msg[len - 1] = 0;
   }
   L.nerr = 0;
   L.errmsg.clear();
   return n;
}

extern "C" int oqh_geometry(const int *Ls, const int *Ltrd)
{
   Lib &L = lib();
   if (!L.inited && oqh_init(0)) return 1;
   for (int mu = 0; mu < 4; mu++)
   {
      if (Ltrd[mu] < 4 || (Ltrd[mu] % 2) || Ls[mu] < 1 || (Ls[mu] % Ltrd[mu]))
      {
         fail("oqh_geometry [global.h]", "improper lattice or thread-block sizes");
         return 1;
      }
   }
   if (L.g.V != 0)
   {
      fail("oqh_geometry", "the geometry can be set only once");
      return 1;
   }
   build_geometry(L.g, Ls, Ltrd);
   size_t V = (size_t)L.g.V;
   L.d_ud = (double2 *)dev_alloc(units_for(L.g.V / 2, 72) * sizeof(double2));
   L.d_u = (float2 *)dev_alloc(units_for(L.g.V / 2, 72) * sizeof(float2));
   L.d_swd = (double2 *)dev_alloc(units_for(L.g.V, 36) * sizeof(double2));
   L.d_sw = (float2 *)dev_alloc(units_for(L.g.V, 36) * sizeof(float2));
   (void)V;
   upload_tables();
   return 0;
}

extern "C" int oqh_volume(void) { return lib().g.V; }

extern "C" void oqh_get_ipt(int *ipt)
{
   Lib &L = lib();
   if (!require_geometry("oqh_get_ipt")) return;
   memcpy(ipt, L.g.ipt.data(), sizeof(int) * (size_t)L.g.V);
}

extern "C" void oqh_set_bc_parms(int type, double cF, double cF_prime)
{
   Lib &L = lib();
   if (type != 0 && type != 3)
   {
      fail("oqh_set_bc_parms [lat_parms.c]",
           "only open (0) and periodic (3) boundary conditions are supported");
      return;
   }
   // lat_parms.c set_bc_parms: cF' enters only for open-SF bc (type 2)
   L.bc = type;
   L.cF[0] = (type < 3) ? cF : 1.0;
   L.cF[1] = (type == 2) ? cF_prime : L.cF[0];
   L.swd_valid = L.sw_valid = false;
   if (L.g.V) upload_tables();
}

extern "C" void oqh_set_sw_parms(double m0, double csw)
{
   Lib &L = lib();
   L.m0 = m0;
   L.csw = csw;
   L.swd_valid = L.sw_valid = false;
}

extern "C" void oqh_set_tm_parms(int eoflg) { lib().eoflg = eoflg != 0; }

extern "C" void oqh_set_sap_parms(const int *bs, int isolv, int nmr, int ncy)
{
   Lib &L = lib();
   if (!require_geometry("oqh_set_sap_parms")) return;
   for (int mu = 0; mu < 4; mu++)
   {
      if (bs[mu] < 4 || (bs[mu] % 2) || (L.g.L[mu] % bs[mu]) || ((L.g.L[mu] / bs[mu]) % 2))
      {
         fail("oqh_set_sap_parms [sap_parms.c]", "improper block sizes");
         return;
      }
   }
   if (isolv < 0 || isolv > 1 || nmr < 1 || ncy < 1)
   {
      fail("oqh_set_sap_parms [sap_parms.c]", "parameters are out of range");
      return;
   }
   memcpy(L.sap_bs, bs, sizeof(L.sap_bs));
   L.sap_isolv = isolv;
   L.sap_nmr = nmr;
   L.sap_ncy = ncy;
}

extern "C" void oqh_get_sap_parms(int *bs, int *isolv, int *nmr, int *ncy)
{
   Lib &L = lib();
   memcpy(bs, L.sap_bs, sizeof(L.sap_bs));
   *isolv = L.sap_isolv;
   *nmr = L.sap_nmr;
   *ncy = L.sap_ncy;
}

extern "C" void oqh_import_ud(const su3_dble *ud)
{
   Lib &L = lib();
   if (!require_geometry("oqh_import_ud")) return;
   int h = L.g.V / 2;
   size_t bytes = (size_t)h * 8 * sizeof(su3_dble);
   void *aos = tmp_buffer(bytes);
   HIP_OK(hipMemcpyAsync(aos, ud, bytes, hipMemcpyHostToDevice, L.stream));
   aos_to_blocks_f64((const double2 *)aos, L.d_ud, h, 72, L.stream);
   HIP_OK(hipStreamSynchronize(L.stream));
   L.ud_set = true;
   L.ud_phase = false;
   L.ud_version++;
   L.u_valid = false;
   L.swd_valid = L.sw_valid = false;
}

extern "C" void oqh_export_ud(su3_dble *ud)
{
   Lib &L = lib();
   if (!require_geometry("oqh_export_ud")) return;
   int h = L.g.V / 2;
   size_t bytes = (size_t)h * 8 * sizeof(su3_dble);
   void *aos = tmp_buffer(bytes);
   blocks_to_aos_f64(L.d_ud, (double2 *)aos, h, 72, L.stream);
   HIP_OK(hipMemcpyAsync(ud, aos, bytes, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}

extern "C" void oqh_export_swd(pauli_dble *swd)
{
   Lib &L = lib();
   if (!require_geometry("oqh_export_swd")) return;
   size_t bytes = (size_t)L.g.V * 2 * sizeof(pauli_dble);
   void *aos = tmp_buffer(bytes);
   blocks_to_aos_f64(L.d_swd, (double2 *)aos, L.g.V, 36, L.stream);
   HIP_OK(hipMemcpyAsync(swd, aos, bytes, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}

extern "C" void oqh_import_swd(const pauli_dble *swd)
{
   Lib &L = lib();
   if (!require_geometry("oqh_import_swd")) return;
   size_t bytes = (size_t)L.g.V * 2 * sizeof(pauli_dble);
   void *aos = tmp_buffer(bytes);
   HIP_OK(hipMemcpyAsync(aos, swd, bytes, hipMemcpyHostToDevice, L.stream));
   aos_to_blocks_f64((const double2 *)aos, L.d_swd, L.g.V, 36, L.stream);
   HIP_OK(hipStreamSynchronize(L.stream));
   L.sws_valid = L.sws32_valid = false;  // an imported term has no S-form
   L.swd_valid = true;
   L.swd_inv[0] = L.swd_inv[1] = false;
   L.sw_valid = false;
   L.swd_epoch++;
}

extern "C" spinor *oqh_alloc_s(int n)
{
   if (!require_geometry("oqh_alloc_s")) return nullptr;
   return (spinor *)register_field(F32, n > 0 ? n : lib().g.V);
}

extern "C" spinor_dble *oqh_alloc_sd(int n)
{
   if (!require_geometry("oqh_alloc_sd")) return nullptr;
   return (spinor_dble *)register_field(F64, n > 0 ? n : lib().g.V);
}

extern "C" void oqh_free(void *p)
{
   Lib &L = lib();
   auto it = L.reg.find((uintptr_t)p);
   if (it == L.reg.end())
   {
      fail("oqh_free", "not a field of this library");
      return;
   }
   HIP_OK(hipStreamSynchronize(L.stream));
   dev_free(p);
   L.reg.erase(it);
}

extern "C" void oqh_import_s(spinor *dev, const spinor *host, int n) { import_field(dev, host, n, F32); }
extern "C" void oqh_export_s(const spinor *dev, spinor *host, int n) { export_field(dev, host, n, F32); }
extern "C" void oqh_import_sd(spinor_dble *dev, const spinor_dble *host, int n) { import_field(dev, host, n, F64); }
extern "C" void oqh_export_sd(const spinor_dble *dev, spinor_dble *host, int n) { export_field(dev, host, n, F64); }

extern "C" void oqh_sync(void)
{
   Lib &L = lib();
   if (L.inited) HIP_OK(hipStreamSynchronize(L.stream));
}

// plain device copy (16-byte accesses, four per lane with the loads issued
// before the stores; non-temporal loads and stores, the fastest of the copy
// forms measured on MI355X) used to measure the achievable HBM bandwidth and
// to calibrate the byte counters of profiles
typedef double v2d_t __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy16(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n)
{
   const v2d_t *av = (const v2d_t *)a;
   v2d_t *bv = (v2d_t *)b;
   size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
   v2d_t v[4];
#pragma unroll
   for (int k = 0; k < 4; k++)
      if (base + 256 * k < n) v[k] = __builtin_nontemporal_load(av + base + 256 * k);
#pragma unroll
   for (int k = 0; k < 4; k++)
      if (base + 256 * k < n) __builtin_nontemporal_store(v[k], bv + base + 256 * k);
}

extern "C" double oqh_copy_bandwidth(double gbytes, int nrep)
{
   Lib &L = lib();
   if (!L.inited && oqh_init(0)) return 0.0;
   size_t n = (size_t)(gbytes * 1e9 / 2.0 / sizeof(double2));
   if (n == 0 || nrep <= 0) return 0.0;
   double2 *a = (double2 *)dev_alloc(n * sizeof(double2));
   double2 *b = (double2 *)dev_alloc(n * sizeof(double2));
   if (!a || !b) return 0.0;
   HIP_OK(hipMemsetAsync(a, 0, n * sizeof(double2), L.stream));
   unsigned nb = (unsigned)((n + 1023) / 1024);
   k_copy16<<<nb, 256, 0, L.stream>>>(a, b, n);
   hipEvent_t e0, e1;
   HIP_OK(hipEventCreate(&e0));
   HIP_OK(hipEventCreate(&e1));
   HIP_OK(hipEventRecord(e0, L.stream));
   for (int k = 0; k < nrep; k++) k_copy16<<<nb, 256, 0, L.stream>>>(a, b, n);
   HIP_OK(hipEventRecord(e1, L.stream));
   HIP_OK(hipEventSynchronize(e1));
   float ms = 0.0f;
   HIP_OK(hipEventElapsedTime(&ms, e0, e1));
   HIP_OK(hipEventDestroy(e0));
   HIP_OK(hipEventDestroy(e1));
   dev_free(a);
   dev_free(b);
   return 2.0 * n * sizeof(double2) * nrep / (ms * 1e-3) / 1e9;
}

// ---------------------------------------------------------------------------
// workspaces (utils/wspace.c)
// ---------------------------------------------------------------------------

static std::vector<int> ws_marks, wsd_marks;

static void clear_ws_marks()
{
   ws_marks.clear();
   wsd_marks.clear();
}

extern "C" void alloc_ws(int n)
{
   Lib &L = lib();
   if (!require_geometry("alloc_ws [wspace.c]")) return;
   while ((int)L.ws.size() < n) L.ws.push_back((spinor *)register_field(F32, L.g.V));
}

extern "C" spinor **reserve_ws(int n)
{
   Lib &L = lib();
   if (n <= 0) return nullptr;
   if (L.ws_used + n > (int)L.ws.size())
   {
      fail("reserve_ws [wspace.c]", "Requested workspace is too large");
      return nullptr;
   }
   spinor **p = (spinor **)malloc(sizeof(spinor *) * n);
   for (int k = 0; k < n; k++)
   {
      p[k] = L.ws[L.ws_used + k];
      HIP_OK(hipMemsetAsync(p[k], 0, units_for(L.g.V, 12) * sizeof(float2), L.stream));
   }
   ws_marks.push_back(n);
   L.ws_used += n;
   L.ws_ptrs.push_back((spinor *)p);
   return p;
}

extern "C" int release_ws(void)
{
   Lib &L = lib();
   if (ws_marks.empty()) return 0;
   int n = ws_marks.back();
   ws_marks.pop_back();
   L.ws_used -= n;
   free(L.ws_ptrs.back());
   L.ws_ptrs.pop_back();
   return n;
}

extern "C" void alloc_wsd(int n)
{
   Lib &L = lib();
   if (!require_geometry("alloc_wsd [wspace.c]")) return;
   while ((int)L.wsd.size() < n) L.wsd.push_back((spinor_dble *)register_field(F64, L.g.V));
}

extern "C" spinor_dble **reserve_wsd(int n)
{
   Lib &L = lib();
   if (n <= 0) return nullptr;
   if (L.wsd_used + n > (int)L.wsd.size())
   {
      fail("reserve_wsd [wspace.c]", "Requested workspace is too large");
      return nullptr;
   }
   spinor_dble **p = (spinor_dble **)malloc(sizeof(spinor_dble *) * n);
   for (int k = 0; k < n; k++)
   {
      p[k] = L.wsd[L.wsd_used + k];
      HIP_OK(hipMemsetAsync(p[k], 0, units_for(L.g.V, 12) * sizeof(double2), L.stream));
   }
   wsd_marks.push_back(n);
   L.wsd_used += n;
   L.wsd_ptrs.push_back((spinor_dble *)p);
   return p;
}

extern "C" int release_wsd(void)
{
   Lib &L = lib();
   if (wsd_marks.empty()) return 0;
   int n = wsd_marks.back();
   wsd_marks.pop_back();
   L.wsd_used -= n;
   free(L.wsd_ptrs.back());
   L.wsd_ptrs.pop_back();
   return n;
}
