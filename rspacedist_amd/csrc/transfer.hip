// transfer.hip — potential assembly on the GPU: the coarse <-> dense grid transfers of
// trans8f.F90 (order-2nf Lagrange prolongation and its adjoint restriction, with the one-rank
// halo handling of overlap_interpolation8f.F90), the multipole halo of trans_d2c_vh, the
// augmented dense density (scf_rhoaugdense8f.F90) and the V_eff sum (scf_potentials8f.F90:
// 255-289).
//
// Every pass is separable and HBM-bound: one thread per output point reads its 2 nf taps along
// one axis (consecutive threads read consecutive x, so the x pass coalesces and the y / z passes
// stream planes), and each output sums its taps in the reference's order (tap ascending, then
// dense point ascending for the restriction), so results match the reference to rounding.
// The restriction is written as a gather (each coarse point collects its 2 nf x nmesh dense
// contributions) instead of the reference's scatter, which needs no atomics.
#include <algorithm>
#include <vector>

#include "kukan_internal.hpp"
#include "transfer.hpp"

namespace kukan {
namespace {

inline unsigned nb_(long n, int t) { return (unsigned)((n + t - 1) / t); }

constexpr int TW_MAX = 2 * 8 * 4;  // 2 nf x nmesh weights, nf <= 8, nmesh <= 4

struct TG {
  int nx, ny, nz, nf, nm, nperi;
  int nxd, nyd, nzd, nf1, nf2, nf3, hm, off;
  long cx, cy, cz;  // coarse halo'd extents (n + 2 nf)
  long dx, dy, dz;  // dense halo'd extents (n nmesh + 2 nf nmesh)
  double aw[TW_MAX];
};

__host__ __device__ inline double AW(const TG &g, int j, int k) { return g.aw[(j + g.nf1) + 2 * g.nf * k]; }
__host__ __device__ inline long CI(const TG &g, int i, int j, int k) {
  return (long)(i + g.nf1) + g.cx * ((long)(j + g.nf1) + g.cy * (long)(k + g.nf1));
}
__host__ __device__ inline long DI(const TG &g, int i, int j, int k) {
  return (long)(i + g.nf3) + g.dx * ((long)(j + g.nf3) + g.dy * (long)(k + g.nf3));
}

// ----------------------------------------------------------------------------- halo handling
// periodic image of a halo'd coordinate (periodic axis), or itself
__device__ inline int wrap_if(int i, int n, bool per) {
  if (!per || (i >= 1 && i <= n)) return i;
  return i < 1 ? i + n : i - n;
}

// out = the one-rank halo fill of overlap_interpolation (nflag = 1): each periodic axis takes
// its halo from the layers it wraps to; an isolated axis keeps in's halo (zero when in is
// given as an interior array, in_halo = 0)
__global__ void k_fill(TG g, const double *in, int in_halo, double *out) {
  const long tot = g.cx * g.cy * g.cz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int i = (int)(q % g.cx) - g.nf1, j = (int)((q / g.cx) % g.cy) - g.nf1, k = (int)(q / (g.cx * g.cy)) - g.nf1;
  const int wi = wrap_if(i, g.nx, g.nperi >= 1), wj = wrap_if(j, g.ny, g.nperi >= 2),
            wk = wrap_if(k, g.nz, g.nperi >= 3);
  double v;
  if (in_halo) {
    v = in[CI(g, wi, wj, wk)];
  } else {
    const bool inside = wi >= 1 && wi <= g.nx && wj >= 1 && wj <= g.ny && wk >= 1 && wk <= g.nz;
    v = inside ? in[(long)(wi - 1) + (long)g.nx * ((long)(wj - 1) + (long)g.ny * (wk - 1))] : 0.0;
  }
  out[q] = v;
}

// overlap_interpolation (nflag = 2) along one periodic axis: per line of the other two axes
// (full extents), halo layer n + iov onto layer iov and layer -nf + iov onto n - nf + iov
__global__ void k_fold(TG g, int axis, double *v) {
  const long m1 = axis == 0 ? g.cy : g.cx, m2 = axis == 2 ? g.cy : g.cz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= m1 * m2) return;
  const int p = (int)(q % m1) - g.nf1, r = (int)(q / m1) - g.nf1;
  const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  auto at = [&](int a) -> double & {
    return axis == 0 ? v[CI(g, a, p, r)] : axis == 1 ? v[CI(g, p, a, r)] : v[CI(g, p, r, a)];
  };
  for (int iov = 1; iov <= g.nf; iov++) {
    at(iov) = at(iov) + at(n + iov);
    at(n - g.nf + iov) = at(n - g.nf + iov) + at(-g.nf + iov);
  }
}

// ----------------------------------------------------------------------------- prolongation
// t1(jx, jy, jz): jx dense interior, jy / jz coarse halo'd
__global__ void k_pro_x(TG g, const double *vc, double *t1) {
  const long tot = (long)g.nxd * g.cy * g.cz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int jx = (int)(q % g.nxd) + 1, jy = (int)((q / g.nxd) % g.cy) - g.nf1, jz = (int)(q / ((long)g.nxd * g.cy)) - g.nf1;
  const int kx = (jx + g.hm) % g.nm, lx = (jx + g.hm) / g.nm;
  double s = 0.0;
  for (int ix = -g.nf1; ix <= g.nf; ix++) s = s + AW(g, ix, kx) * vc[CI(g, lx + ix, jy, jz)];
  t1[q] = s;
}
// t2(jx, jy, jz): jx, jy dense interior, jz coarse halo'd
__global__ void k_pro_y(TG g, const double *t1, double *t2) {
  const long tot = (long)g.nxd * g.nyd * g.cz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int jx = (int)(q % g.nxd), jy = (int)((q / g.nxd) % g.nyd) + 1, jz = (int)(q / ((long)g.nxd * g.nyd));
  const int ky = (jy + g.hm) % g.nm, ly = (jy + g.hm) / g.nm;
  double s = 0.0;
  for (int iy = -g.nf1; iy <= g.nf; iy++)
    s = s + AW(g, iy, ky) * t1[jx + (long)g.nxd * ((long)(ly + iy + g.nf1) + g.cy * jz)];
  t2[q] = s;
}
// dense interior (ld = interior array, or the halo'd dense array when halo = 1)
__global__ void k_pro_z(TG g, const double *t2, double *out, int halo) {
  const long tot = (long)g.nxd * g.nyd * g.nzd;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int jx = (int)(q % g.nxd), jy = (int)((q / g.nxd) % g.nyd), jz = (int)(q / ((long)g.nxd * g.nyd)) + 1;
  const int kz = (jz + g.hm) % g.nm, lz = (jz + g.hm) / g.nm;
  double s = 0.0;
  for (int iz = -g.nf1; iz <= g.nf; iz++)
    s = s + AW(g, iz, kz) * t2[jx + (long)g.nxd * (jy + (long)g.nyd * (long)(lz + iz + g.nf1))];
  out[halo ? DI(g, jx + 1, jy + 1, jz) : q] = s;
}

// ----------------------------------------------------------------------------- restriction
// the halo of the dense array along the periodic axes is cleared (interpolation8f.F90:126-177)
__global__ void k_zero_periodic_halo(TG g, double *vd) {
  const long tot = g.dx * g.dy * g.dz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int i = (int)(q % g.dx) - g.nf3, j = (int)((q / g.dx) % g.dy) - g.nf3, k = (int)(q / (g.dx * g.dy)) - g.nf3;
  const bool hx = i < 1 || i > g.nxd, hy = j < 1 || j > g.nyd, hz = k < 1 || k > g.nzd;
  if ((hx && g.nperi > 0) || (hy && g.nperi > 1) || (hz && g.nperi > 2)) vd[q] = 0.0;
}
// coarse index c along an axis collects, tap i ascending then dense point ascending, the dense
// points j with (j + off) / nmesh - nf + i == c, weight aweight(i, (j + off) % nmesh)
// t2r(jx, jy, cz): jx, jy dense halo'd, cz coarse halo'd
__global__ void k_res_z(TG g, const double *vd, double *t2r) {
  const long tot = g.dx * g.dy * g.cz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const long jxy = q % (g.dx * g.dy);
  const int cz = (int)(q / (g.dx * g.dy)) - g.nf1;
  double s = 0.0;
  for (int iz = -g.nf1; iz <= g.nf; iz++) {
    const int j0 = (cz - iz + g.nf) * g.nm - g.off;
    for (int t = 0; t < g.nm; t++) {
      const int jz = j0 + t;
      if (jz < -g.nf3 || jz > g.nzd + g.nf2) continue;
      s = s + AW(g, iz, t) * vd[jxy + g.dx * g.dy * (long)(jz + g.nf3)];
    }
  }
  t2r[q] = s;
}
// t1r(jx, cy, cz): jx dense halo'd, cy / cz coarse halo'd
__global__ void k_res_y(TG g, const double *t2r, double *t1r) {
  const long tot = g.dx * g.cy * g.cz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const long jx = q % g.dx;
  const int cy = (int)((q / g.dx) % g.cy) - g.nf1;
  const long cz = q / (g.dx * g.cy);
  double s = 0.0;
  for (int iy = -g.nf1; iy <= g.nf; iy++) {
    const int j0 = (cy - iy + g.nf) * g.nm - g.off;
    for (int t = 0; t < g.nm; t++) {
      const int jy = j0 + t;
      if (jy < -g.nf3 || jy > g.nyd + g.nf2) continue;
      s = s + AW(g, iy, t) * t2r[jx + g.dx * ((long)(jy + g.nf3) + g.dy * cz)];
    }
  }
  t1r[q] = s;
}
// vc(cx, cy, cz) coarse halo'd, divided by nmesh^3
__global__ void k_res_x(TG g, const double *t1r, double *vc) {
  const long tot = g.cx * g.cy * g.cz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int cx = (int)(q % g.cx) - g.nf1;
  const long cyz = q / g.cx;
  double s = 0.0;
  for (int ix = -g.nf1; ix <= g.nf; ix++) {
    const int j0 = (cx - ix + g.nf) * g.nm - g.off;
    for (int t = 0; t < g.nm; t++) {
      const int jx = j0 + t;
      if (jx < -g.nf3 || jx > g.nxd + g.nf2) continue;
      s = s + AW(g, ix, t) * t1r[(long)(jx + g.nf3) + g.dx * cyz];
    }
  }
  vc[q] = s / (double)(g.nm * g.nm * g.nm);
}

// ----------------------------------------------------------------------------- dense halos
// the potential of every atom's point multipoles added once per isolated face a dense halo
// point lies in, atom by atom (trans8f.F90:155-412, one rank)
__global__ void k_multipole_halo(TG g, int natom, double ddx, double ddy, double ddz, double xmax,
                                 double ymax, double zmax, const double *mp, const double *cp1,
                                 const double *atx, const double *aty, const double *atz, double *vd) {
  const long tot = g.dx * g.dy * g.dz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int jx = (int)(q % g.dx) - g.nf3, jy = (int)((q / g.dx) % g.dy) - g.nf3, jz = (int)(q / (g.dx * g.dy)) - g.nf3;
  const bool fx = g.nperi == 0 && (jx < 1 || jx > g.nxd);
  const bool fy = g.nperi <= 1 && (jy < 1 || jy > g.nyd);
  const bool fz = g.nperi <= 2 && (jz < 1 || jz > g.nzd);
  if (!fx && !fy && !fz) return;
  double v = vd[q];
  for (int na = 0; na < natom; na++) {
    const double c1 = cp1[na];
    const double *m = mp + 10 * (long)na;
    const double atptmp = m[0] / c1, edpolx = m[1] / c1, edpoly = m[2] / c1, edpolz = m[3] / c1;
    const double eqpoxx = m[4] / c1, eqpoyy = m[5] / c1, eqpozz = m[6] / c1;
    const double eqpoxy2 = 2.0 * (m[7] / c1), eqpoyz2 = 2.0 * (m[8] / c1), eqpozx2 = 2.0 * (m[9] / c1);
    const double x = atx[na] - (jx * ddx - xmax - 0.5 * ddx);
    const double y = aty[na] - (jy * ddy - ymax - 0.5 * ddy);
    const double z = atz[na] - (jz * ddz - zmax - 0.5 * ddz);
    const double r = sqrt(x * x + y * y + z * z);
    const double rin = 1.0 / r, r2in = rin * rin, r3in = r2in * rin, r5in = r3in * r2in;
    const double vmopm = -c1 * rin, derfdrr = -r3in;
    const double vdpmx = c1 * derfdrr * x, vdpmy = c1 * derfdrr * y, vdpmz = c1 * derfdrr * z;
    const double vqmxx = -c1 * (3.0 * x * x * r5in - r3in), vqmyy = -c1 * (3.0 * y * y * r5in - r3in);
    const double vqmzz = -c1 * (3.0 * z * z * r5in - r3in), vqmxy = -c1 * 3.0 * x * y * r5in;
    const double vqmyz = -c1 * 3.0 * y * z * r5in, vqmzx = -c1 * 3.0 * z * x * r5in;
    const int times = (int)fx + (int)fy + (int)fz;  // once per face loop of the reference
    for (int t = 0; t < times; t++)
      v = v - atptmp * vmopm - edpolx * vdpmx - edpoly * vdpmy - edpolz * vdpmz - eqpoxx * vqmxx -
          eqpoyy * vqmyy - eqpozz * vqmzz - eqpoxy2 * vqmxy - eqpoyz2 * vqmyz - eqpozx2 * vqmzx;
  }
  vd[q] = v;
}
// the constant of trans_d2c_vxc on the isolated faces of the dense halo
__global__ void k_const_halo(TG g, double c, double *vd) {
  const long tot = g.dx * g.dy * g.dz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int i = (int)(q % g.dx) - g.nf3, j = (int)((q / g.dx) % g.dy) - g.nf3, k = (int)(q / (g.dx * g.dy)) - g.nf3;
  const bool hx = i < 1 || i > g.nxd, hy = j < 1 || j > g.nyd, hz = k < 1 || k > g.nzd;
  if ((hx && g.nperi == 0) || (hy && g.nperi <= 1) || (hz && g.nperi <= 2)) vd[q] = c;
}
// interior <-> halo'd copies
__global__ void k_dense_in(TG g, const double *src, double *vd) {
  const long tot = (long)g.nxd * g.nyd * g.nzd;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int i = (int)(q % g.nxd) + 1, j = (int)((q / g.nxd) % g.nyd) + 1, k = (int)(q / ((long)g.nxd * g.nyd)) + 1;
  vd[DI(g, i, j, k)] = src[q];
}
__global__ void k_coarse_out(TG g, const double *vc, double *dst) {
  const long tot = (long)g.nx * g.ny * g.nz;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int i = (int)(q % g.nx) + 1, j = (int)((q / g.nx) % g.ny) + 1, k = (int)(q / ((long)g.nx * g.ny)) + 1;
  dst[q] = vc[CI(g, i, j, k)];
}

// ----------------------------------------------------------------------------- density, V_eff
__global__ void k_spin_sum(long n, int nspv, const double *rho, double *out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = rho[i];
  if (nspv > 1) v = v + rho[i + n];
  out[i] = v;
}
// rows (dense points) with augmentation entries, each summed in the reference's atom / list order
__global__ void k_aug_rows(long nrows, const int *rows, const int *row_ptr, const int *ent,
                           const double *aug, double *out) {
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int p = rows[r];
  double v = out[p];
  for (int e = row_ptr[r]; e < row_ptr[r + 1]; e++) v = v + aug[ent[e]];
  out[p] = v;
}
__global__ void k_veff(int nx, int ny, int nz, int nspv, double xmax, double ymax, double zmax,
                       double dx, double dy, double dz, double bx, double by, double bz,
                       const double *vloc, const double *vh, const double *vx, double *veff) {
  const long n = (long)nx * ny * nz;
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ix = (int)(p % nx) + 1, iy = (int)((p / nx) % ny) + 1, iz = (int)(p / ((long)nx * ny)) + 1;
  const double vdia = vloc[p] + vh[p] + bx * (((double)ix - 0.5) * dx - xmax) +
                      by * (((double)iy - 0.5) * dy - ymax) + bz * (((double)iz - 0.5) * dz - zmax);
  for (int ns = 0; ns < nspv; ns++) {
    double v = vx[p + n * ns];
    if (nspv < 3 || ns < 2) v = v + vdia;
    veff[p + n * ns] = v;
  }
}

#define LAUNCH(kern, count, ...)                                                     \
  do {                                                                               \
    if ((count) > 0) hipLaunchKernelGGL(kern, dim3(nb_((count), 256)), dim3(256), 0, s, __VA_ARGS__); \
    HIPCK(hipGetLastError());                                                        \
  } while (0)

TG make_tg(int nx, int ny, int nz, int nf, int nmesh, int nperi) {
  if (nf < 1 || nf > 8) throw Error(1, "interpolation: nf must be in 1-8");
  if (nmesh < 1 || nmesh > 4) throw Error(1, "interpolation: nmesh must be in 1-4");
  if (nperi >= 1 && nf > nx) throw Error(1, "interpolation: the periodic halo is wider than the cell");
  if (nperi >= 2 && nf > ny) throw Error(1, "interpolation: the periodic halo is wider than the cell");
  if (nperi >= 3 && nf > nz) throw Error(1, "interpolation: the periodic halo is wider than the cell");
  TG g{};
  g.nx = nx; g.ny = ny; g.nz = nz; g.nf = nf; g.nm = nmesh; g.nperi = nperi;
  g.nxd = nx * nmesh; g.nyd = ny * nmesh; g.nzd = nz * nmesh;
  g.nf1 = nf - 1; g.nf2 = nf * nmesh; g.nf3 = g.nf2 - 1; g.hm = (nmesh - 1) / 2;
  g.off = nf * nmesh + g.hm;
  g.cx = nx + 2 * nf; g.cy = ny + 2 * nf; g.cz = nz + 2 * nf;
  g.dx = g.nxd + 2L * g.nf2; g.dy = g.nyd + 2L * g.nf2; g.dz = g.nzd + 2L * g.nf2;
  // Lagrange weights of the 2 nf taps for the nmesh dense offsets (interpolation8f.F90:40-50)
  const double tmp = 1.0 / nmesh, tmp1 = (nmesh % 2 == 1) ? 0.0 : 0.5;
  for (int k = 0; k < nmesh; k++) {
    const double r = ((double)k + tmp1) * tmp;
    for (int j = -g.nf1; j <= nf; j++) {
      double a = 1.0;
      for (int i = -g.nf1; i <= nf; i++)
        if (i != j) a = a * (r - (double)i) / ((double)j - (double)i);
      g.aw[(j + g.nf1) + 2 * nf * k] = a;
    }
  }
  return g;
}

}  // namespace

// ============================================================================= TransferGrid
struct TransferGrid::Impl {
  TG g;
};
TransferGrid::TransferGrid(int nx, int ny, int nz, int nf, int nmesh, int nperi)
    : impl(new Impl{make_tg(nx, ny, nz, nf, nmesh, nperi)}) {}
TransferGrid::~TransferGrid() { delete impl; }
long TransferGrid::coarse_halo_count() const { return impl->g.cx * impl->g.cy * impl->g.cz; }
long TransferGrid::dense_halo_count() const { return impl->g.dx * impl->g.dy * impl->g.dz; }
long TransferGrid::dense_count() const { return (long)impl->g.nxd * impl->g.nyd * impl->g.nzd; }
long TransferGrid::coarse_count() const { return (long)impl->g.nx * impl->g.ny * impl->g.nz; }

void TransferGrid::fill(const double *in, bool in_halo, double *out_halo, hipStream_t s) const {
  const TG &g = impl->g;
  LAUNCH(k_fill, g.cx * g.cy * g.cz, g, in, in_halo ? 1 : 0, out_halo);
}

void TransferGrid::fold(double *vc, hipStream_t s) const {
  const TG &g = impl->g;
  for (int axis = 0; axis < 3; axis++) {
    if (g.nperi < axis + 1) continue;
    const long lines = (axis == 0 ? g.cy : g.cx) * (axis == 2 ? g.cy : g.cz);
    LAUNCH(k_fold, lines, g, axis, vc);
  }
}

void TransferGrid::prolong(const double *vc_halo, double *out, bool out_halo, double *t1, double *t2,
                           hipStream_t s) const {
  const TG &g = impl->g;
  if (out_halo) HIPCK(hipMemsetAsync(out, 0, sizeof(double) * dense_halo_count(), s));
  LAUNCH(k_pro_x, (long)g.nxd * g.cy * g.cz, g, vc_halo, t1);
  LAUNCH(k_pro_y, (long)g.nxd * g.nyd * g.cz, g, t1, t2);
  LAUNCH(k_pro_z, (long)g.nxd * g.nyd * g.nzd, g, t2, out, out_halo ? 1 : 0);
}

void TransferGrid::restrict_(double *vd_halo, double *vc_halo, double *t2r, double *t1r,
                             hipStream_t s) const {
  const TG &g = impl->g;
  if (g.nperi > 0) LAUNCH(k_zero_periodic_halo, g.dx * g.dy * g.dz, g, vd_halo);
  LAUNCH(k_res_z, g.dx * g.dy * g.cz, g, vd_halo, t2r);
  LAUNCH(k_res_y, g.dx * g.cy * g.cz, g, t2r, t1r);
  LAUNCH(k_res_x, g.cx * g.cy * g.cz, g, t1r, vc_halo);
}

void TransferGrid::dense_in(const double *src, double *vd_halo, hipStream_t s) const {
  const TG &g = impl->g;
  LAUNCH(k_dense_in, (long)g.nxd * g.nyd * g.nzd, g, src, vd_halo);
}
void TransferGrid::coarse_out(const double *vc_halo, double *dst, hipStream_t s) const {
  const TG &g = impl->g;
  LAUNCH(k_coarse_out, (long)g.nx * g.ny * g.nz, g, vc_halo, dst);
}
void TransferGrid::multipole_halo(int natom, double xmax, double ymax, double zmax,
                                  const double *atmpole, const double *cp1, const double *atx,
                                  const double *aty, const double *atz, double *vd_halo,
                                  hipStream_t s) const {
  const TG &g = impl->g;
  if (g.nperi >= 3 || natom <= 0) return;
  const double ddx = 2.0 * xmax / g.nxd, ddy = 2.0 * ymax / g.nyd, ddz = 2.0 * zmax / g.nzd;
  LAUNCH(k_multipole_halo, g.dx * g.dy * g.dz, g, natom, ddx, ddy, ddz, xmax, ymax, zmax, atmpole, cp1,
         atx, aty, atz, vd_halo);
}
void TransferGrid::const_halo(double c, double *vd_halo, hipStream_t s) const {
  const TG &g = impl->g;
  if (g.nperi >= 3) return;
  LAUNCH(k_const_halo, g.dx * g.dy * g.dz, g, c, vd_halo);
}

void spin_sum(long n, int nspv, const double *rho, double *out, hipStream_t s) {
  LAUNCH(k_spin_sum, n, n, nspv, rho, out);
}

void AugPlan::build(int key_natpri_inps, int key_pp_paw, int natom, long nd, int num_list_d,
                    const std::vector<int> &indspe, const std::vector<int> &natprid,
                    const std::vector<int> &napsd, const std::vector<int> &natinfd,
                    const std::vector<int> &ntyppp, const std::vector<int> &lstvecd2) {
  // rows in ascending point order, each row's entries in the reference's atom / list order
  std::vector<int> cnt(nd + 1, 0);
  std::vector<std::pair<int, int>> order;  // (point, entry index into rhoaug3d)
  for (int na = 0; na < natom; na++) {
    if (indspe[na] < 1 || indspe[na] > (int)ntyppp.size())
      throw Error(1, "scf_rhoaugdense: indspe out of range");
    if (ntyppp[indspe[na] - 1] != key_pp_paw || natprid[na] != key_natpri_inps) continue;
    if (napsd[na] < 1 || (long)napsd[na] * num_list_d > (long)lstvecd2.size() || natinfd[na] > num_list_d)
      throw Error(1, "scf_rhoaugdense: napsd / natinfd outside the list table");
    const long base = (long)(napsd[na] - 1) * num_list_d;
    for (int i = 0; i < natinfd[na]; i++) {
      const int p = lstvecd2[base + i] - 1;
      if (p < 0 || p >= nd) throw Error(1, "scf_rhoaugdense: list entry outside the dense grid");
      order.emplace_back(p, (int)(base + i));
      cnt[p + 1]++;
    }
  }
  std::vector<int> rptr, rws, entries(order.size());
  for (long p = 0; p < nd; p++) cnt[p + 1] += cnt[p];
  std::vector<int> fillpos(cnt.begin(), cnt.end() - 1);
  for (const auto &o : order) entries[fillpos[o.first]++] = o.second;  // stable: keeps the order
  for (long p = 0; p < nd; p++)
    if (cnt[p + 1] > cnt[p]) {
      rws.push_back((int)p);
      rptr.push_back(cnt[p]);
    }
  rptr.push_back((int)entries.size());
  nrows = (long)rws.size();
  rows_h = rws;
  row_ptr_h = rptr;
  ent_h = entries;
}

void aug_add(const AugPlan &pl, const int *rows, const int *row_ptr, const int *ent, const double *aug,
             double *out, hipStream_t s) {
  LAUNCH(k_aug_rows, pl.nrows, pl.nrows, rows, row_ptr, ent, aug, out);
}

void veff_sum(int nx, int ny, int nz, int nspv, double xmax, double ymax, double zmax, double dx,
              double dy, double dz, double bx, double by, double bz, const double *vloc,
              const double *vh, const double *vx, double *veff, hipStream_t s) {
  LAUNCH(k_veff, (long)nx * ny * nz, nx, ny, nz, nspv, xmax, ymax, zmax, dx, dy, dz, bx, by, bz, vloc,
         vh, vx, veff);
}

}  // namespace kukan
