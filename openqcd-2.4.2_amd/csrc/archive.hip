// archive.hip -- exchange of gauge-field configurations and of
// double-precision quark fields in openQCD's "universal" file format
// (archive/archive.c:769-977, archive/sarchive.c:424-580,
// main/README.io:39-85).
//
// File layout (little endian): the lattice sizes N0..N3 as 4-byte integers,
// the average plaquette plaq_sum_dble()/(6 N0 N1 N2 N3) as a double, then,
// for the odd points x in lexicographic order (x3 fastest), the 8 links
// U(x,+0), U(x-0,0), ..., U(x,+3), U(x-3,3) as 18 doubles each.  With the
// whole lattice in one process the records of a point are the 8 consecutive
// entries of udfld() at 8*(ipt[x]-V/2).
//
// Quark-field files: the sizes N0..N3 and eo (0: all points, 1: the even
// points only) as 4-byte integers, the square norm norm_square_dble() of the
// written points as a double, then the spinors (24 doubles) of the written
// points in lexicographic order.
//
// The fields move through host memory in the reference's layouts; the
// plaquette and the square norm are computed on the device.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "oqh_internal.h"

extern "C" double plaq_sum_dble(int icom);
extern "C" qflt norm_square_dble(int vol, int icom, spinor_dble *sd);

namespace oqh {
namespace {

bool host_is_big_endian()
{
   const uint32_t one = 1;
   unsigned char b;
   memcpy(&b, &one, 1);
   return b == 0;
}

void swap_bytes(void *p, size_t size, size_t n)
{
   unsigned char *c = (unsigned char *)p;
   for (size_t i = 0; i < n; i++, c += size)
      for (size_t k = 0; k < size / 2; k++)
      {
         unsigned char t = c[k];
         c[k] = c[size - 1 - k];
         c[size - 1 - k] = t;
      }
}

int lex_index(const int *L, const int *x) { return x[3] + L[3] * (x[2] + L[2] * (x[1] + L[1] * x[0])); }

bool is_zero(const su3_dble *u)
{
   const double *r = (const double *)u;
   for (int i = 0; i < 18; i++)
      if (r[i] != 0.0) return false;
   return true;
}

// lattice/bcnds.c:409-457 (boundary types 0 and 3): with open bc the links
// U(x,-0) at x0 = 0 and U(x,+0) at x0 = N0-1 vanish, every other link is
// nonzero
bool check_zero(const Lib &L, const std::vector<su3_dble> &ud)
{
   const int h = L.g.V / 2, N0 = L.g.L[0];
   for (int j = 0; j < h; j++)
   {
      const su3_dble *u = ud.data() + 8 * (size_t)j;
      const int t = L.g.x0[h + j];
      for (int l = 0; l < 8; l++)
      {
         bool zero_expected = L.bc == 0 && ((t == 0 && l == 1) || (t == N0 - 1 && l == 0));
         if (is_zero(u + l) != zero_expected) return false;
      }
   }
   return true;
}

}  // namespace
}  // namespace oqh

using namespace oqh;

extern "C" void export_cnfg(char *out)
{
   Lib &L = lib();
   const char *where = "export_cnfg [archive.c]";
   if (!require_geometry(where)) return;
   if (!L.ud_set)
   {
      fail(where, "gauge field is not set");
      return;
   }
   if (L.ud_phase)
   {
      fail(where, "Attempt to export phase-modified gauge field");
      return;
   }
   const int *N = L.g.L, h = L.g.V / 2;
   std::vector<su3_dble> ud(8 * (size_t)h);
   oqh_export_ud(ud.data());
   double nplaq = (double)(6 * N[0] * N[1]) * (double)(N[2] * N[3]);
   double plaq = plaq_sum_dble(1) / nplaq;

   FILE *f = fopen(out, "wb");
   if (f == nullptr)
   {
      fail(where, "Unable to open output file");
      return;
   }
   const bool swap = host_is_big_endian();
   int32_t lsize[4] = {N[0], N[1], N[2], N[3]};
   if (swap)
   {
      swap_bytes(lsize, 4, 4);
      swap_bytes(&plaq, 8, 1);
   }
   size_t iw = fwrite(lsize, sizeof(int32_t), 4, f);
   iw += fwrite(&plaq, sizeof(double), 1, f);
   bool bad = iw != 5;
   su3_dble rec[8];
   int x[4];
   for (x[0] = 0; x[0] < N[0] && !bad; x[0]++)
      for (x[1] = 0; x[1] < N[1]; x[1]++)
         for (x[2] = 0; x[2] < N[2]; x[2]++)
            for (x[3] = (x[0] + x[1] + x[2] + 1) & 1; x[3] < N[3]; x[3] += 2)
            {
               const int ix = L.g.ipt[lex_index(N, x)];
               memcpy(rec, ud.data() + 8 * (size_t)(ix - h), sizeof(rec));
               if (swap) swap_bytes(rec, 8, 8 * 18);
               bad |= fwrite(rec, sizeof(su3_dble), 8, f) != 8;
            }
   fclose(f);
   if (bad) fail(where, "Incorrect write count");
}

// archive.c:769-800
extern "C" void lat_sizes(char *in, int *nl)
{
   const char *where = "lat_sizes [archive.c]";
   FILE *f = fopen(in, "rb");
   if (f == nullptr)
   {
      fail(where, "Unable to open input file");
      return;
   }
   int32_t lsize[4];
   size_t ir = fread(lsize, sizeof(int32_t), 4, f);
   fclose(f);
   if (ir != 4)
   {
      fail(where, "Incorrect read count");
      return;
   }
   if (host_is_big_endian()) swap_bytes(lsize, 4, 4);
   for (int mu = 0; mu < 4; mu++) nl[mu] = (int)lsize[mu];
}

// archive.c:873-977.  The whole lattice lives in one process, so the sizes
// on the file must equal the current ones (the extension of smaller fields
// needs sizes that are multiples of the local lattice, which here is the
// whole lattice); mask is checked as the reference does.
extern "C" void import_cnfg(char *in, int mask)
{
   Lib &L = lib();
   const char *where = "import_cnfg [archive.c]";
   if (!require_geometry(where)) return;
   FILE *f = fopen(in, "rb");
   if (f == nullptr)
   {
      fail(where, "Unable to open input file");
      return;
   }
   const bool swap = host_is_big_endian();
   int32_t lsize[4];
   double plaq0;
   size_t ir = fread(lsize, sizeof(int32_t), 4, f);
   ir += fread(&plaq0, sizeof(double), 1, f);
   if (ir != 5)
   {
      fclose(f);
      fail(where, "Incorrect read count");
      return;
   }
   if (swap)
   {
      swap_bytes(lsize, 4, 4);
      swap_bytes(&plaq0, 8, 1);
   }
   // check_lattice_size (archive.c:335-384) with local sizes = global sizes
   const int *N = L.g.L;
   int ie = 0;
   for (int mu = 0; mu < 4; mu++)
   {
      const int nl = (int)lsize[mu];
      if (nl < 1 || (N[mu] % nl)) ie |= 0x1;
      if (nl < 1 || (nl % N[mu])) ie |= 0x2;
      if (nl >= 1 && ((mask >> mu) & 0x1) && nl != N[mu] && ((N[mu] / nl) & 0x1)) ie |= 0x4;
   }
   const char *msg = nullptr;
   if (ie & 0x1)
      msg = "Previous lattice sizes do not divide the current ones";
   else if (ie & 0x2)
      msg = "Previous lattice sizes are not multiples of the local ones";
   else if (ie & 0x4)
      msg = "Unsuitable lattice sizes for anti-periodic extension";
   if (msg)
   {
      fclose(f);
      fail("check_lattice_size [archive.c]", msg);
      return;
   }

   const int h = L.g.V / 2;
   std::vector<su3_dble> ud(8 * (size_t)h);
   bool bad = false;
   int x[4];
   for (x[0] = 0; x[0] < N[0] && !bad; x[0]++)
      for (x[1] = 0; x[1] < N[1]; x[1]++)
         for (x[2] = 0; x[2] < N[2]; x[2]++)
            for (x[3] = (x[0] + x[1] + x[2] + 1) & 1; x[3] < N[3]; x[3] += 2)
            {
               const int ix = L.g.ipt[lex_index(N, x)];
               su3_dble *rec = ud.data() + 8 * (size_t)(ix - h);
               bad |= fread(rec, sizeof(su3_dble), 8, f) != 8;
               if (swap) swap_bytes(rec, 8, 8 * 18);
            }
   fclose(f);
   if (bad)
   {
      fail(where, "Incorrect read count");
      return;
   }
   oqh_import_ud(ud.data());
   if (!check_zero(L, ud))
   {
      fail(where, "Incompatible boundary conditions");
      return;
   }
   double nplaq = (double)(6 * N[0] * N[1]) * (double)(N[2] * N[3]);
   plaq0 *= nplaq;
   double plaq1 = plaq_sum_dble(1);
   double eps = sqrt(64.0 * nplaq) * DBL_EPSILON;
   eps += 64.0 * (fabs(plaq0) + fabs(plaq1)) * DBL_EPSILON;
   if (fabs(plaq1 - plaq0) > eps) fail(where, "Incorrect average plaquette");
}

// sarchive.c:424-503.  sd is a device field of the library or a host array
// in the layout of the reference (ipt order, even points first).
extern "C" void export_sfld(char *out, int eo, spinor_dble *sd)
{
   Lib &L = lib();
   const char *where = "export_sfld [sarchive.c]";
   if (!require_geometry(where)) return;
   if (eo < 0 || eo > 1)
   {
      fail(where, "Parameter eo is out of range");
      return;
   }
   if (sd == nullptr)
   {
      fail(where, "Attempt to access unallocated memory space");
      return;
   }
   const int *N = L.g.L, V = L.g.V, n = eo ? V / 2 : V;
   const int vt = L.g.VT;
   double norm = norm_square_dble(eo ? vt / 2 : vt, 3, sd).q[0];
   std::vector<spinor_dble> buf;
   const spinor_dble *src = sd;
   int off = 0;
   if (device_field(sd, F64, &off))
   {
      buf.resize((size_t)n);
      oqh_export_sd(sd, buf.data(), n);
      src = buf.data();
   }
   FILE *f = fopen(out, "wb");
   if (f == nullptr)
   {
      fail(where, "Unable to open output file");
      return;
   }
   const bool swap = host_is_big_endian();
   int32_t lsize[5] = {N[0], N[1], N[2], N[3], eo};
   if (swap)
   {
      swap_bytes(lsize, 4, 5);
      swap_bytes(&norm, 8, 1);
   }
   size_t iw = fwrite(lsize, sizeof(int32_t), 5, f);
   iw += fwrite(&norm, sizeof(double), 1, f);
   bool bad = iw != 6;
   std::vector<spinor_dble> line((size_t)N[3]);
   int x[4];
   for (x[0] = 0; x[0] < N[0] && !bad; x[0]++)
      for (x[1] = 0; x[1] < N[1]; x[1]++)
         for (x[2] = 0; x[2] < N[2]; x[2]++)
         {
            int m = 0;
            for (x[3] = eo ? (x[0] + x[1] + x[2]) & 1 : 0; x[3] < N[3]; x[3] += eo ? 2 : 1)
               line[m++] = src[L.g.ipt[lex_index(N, x)]];
            if (swap) swap_bytes(line.data(), 8, (size_t)m * 24);
            bad |= fwrite(line.data(), sizeof(spinor_dble), (size_t)m, f) != (size_t)m;
         }
   fclose(f);
   if (bad) fail(where, "Incorrect write count");
}

// sarchive.c:522-580
extern "C" void import_sfld(char *in, int eo, spinor_dble *sd)
{
   Lib &L = lib();
   const char *where = "import_sfld [sarchive.c]";
   if (!require_geometry(where)) return;
   if (eo < 0 || eo > 1)
   {
      fail(where, "Parameter eo is out of range");
      return;
   }
   if (sd == nullptr)
   {
      fail(where, "Attempt to access unallocated memory space");
      return;
   }
   const int *N = L.g.L, V = L.g.V, n = eo ? V / 2 : V;
   const int vt = L.g.VT;
   FILE *f = fopen(in, "rb");
   if (f == nullptr)
   {
      fail(where, "Unable to open input file");
      return;
   }
   const bool swap = host_is_big_endian();
   int32_t lsize[5];
   double norm0;
   size_t ir = fread(lsize, sizeof(int32_t), 5, f);
   ir += fread(&norm0, sizeof(double), 1, f);
   if (ir != 6)
   {
      fclose(f);
      fail(where, "Incorrect read count");
      return;
   }
   if (swap)
   {
      swap_bytes(lsize, 4, 5);
      swap_bytes(&norm0, 8, 1);
   }
   if (lsize[0] != N[0] || lsize[1] != N[1] || lsize[2] != N[2] || lsize[3] != N[3] || lsize[4] != eo)
   {
      fclose(f);
      fail(where, "Lattice sizes do not match");
      return;
   }
   int off = 0;
   const bool dev = device_field(sd, F64, &off) != nullptr;
   std::vector<spinor_dble> buf;
   spinor_dble *dst = sd;
   if (dev)
   {
      buf.resize((size_t)n);
      dst = buf.data();
   }
   std::vector<spinor_dble> line((size_t)N[3]);
   bool bad = false;
   int x[4];
   for (x[0] = 0; x[0] < N[0] && !bad; x[0]++)
      for (x[1] = 0; x[1] < N[1]; x[1]++)
         for (x[2] = 0; x[2] < N[2]; x[2]++)
         {
            const int x3a = eo ? (x[0] + x[1] + x[2]) & 1 : 0, m = eo ? N[3] / 2 : N[3];
            bad |= fread(line.data(), sizeof(spinor_dble), (size_t)m, f) != (size_t)m;
            if (swap) swap_bytes(line.data(), 8, (size_t)m * 24);
            int k = 0;
            for (x[3] = x3a; x[3] < N[3]; x[3] += eo ? 2 : 1) dst[L.g.ipt[lex_index(N, x)]] = line[k++];
         }
   fclose(f);
   if (bad)
   {
      fail(where, "Incorrect read count");
      return;
   }
   if (dev) oqh_import_sd(sd, buf.data(), n);
   double norm1 = norm_square_dble(eo ? vt / 2 : vt, 3, sd).q[0];
   double eps = 64.0 * (norm0 + norm1) * DBL_EPSILON;
   if (fabs(norm1 - norm0) > eps) fail(where, "Incorrect square norm");
}
