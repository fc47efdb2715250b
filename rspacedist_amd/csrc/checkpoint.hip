// checkpoint.hip — the checkpoint files of rw_wfpot8f.F90 (wfc.* version 5, rho.* / r_o.* /
// vht.* version 3): Fortran sequential unformatted records (each framed by its 4-byte byte
// count), one rank.  Arrays may be host or device pointers.  Device arrays never go through a
// whole host copy: records stream through two pinned staging buffers, the device-to-host copy
// of one chunk overlapping the file write of the previous one (and the file read of one chunk
// the host-to-device copy of the previous one on the way in).
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "abi_common.hpp"
#include "kukan_hip.h"

using namespace kukan;
using namespace kukan::abi;

namespace {

constexpr int WFC_VERSION = 5, RHO_VERSION = 3, VHT_VERSION = 3;
constexpr size_t STAGE = 32u << 20;  // bytes per staging buffer

// rw_wfpot8f.F90:535-572 get_fname: chdir [/] chnam.nnnnn[.kkk]
std::string ckpt_path(const char *chdir, const char *chnam, int nr1, int nr2) {
  if (nr1 >= 100000 || (nr2 >= 0 && nr2 >= 1000)) throw Error(1, "get_fname: not enough digits");
  std::string cd = chdir ? chdir : "";
  while (!cd.empty() && cd.back() == ' ') cd.pop_back();
  char tail[64];
  if (nr2 >= 0) snprintf(tail, sizeof tail, "%s.%05d.%03d", chnam, nr1, nr2);
  else snprintf(tail, sizeof tail, "%s.%05d", chnam, nr1);
  return cd.empty() ? std::string(tail) : cd + "/" + tail;
}

struct File {
  FILE *f = nullptr;
  File(const std::string &p, const char *mode) { f = fopen(p.c_str(), mode); }
  ~File() {
    if (f) fclose(f);
  }
};

// two pinned buffers, allocated once per process
struct Stage {
  char *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  Stage() {
    for (int i = 0; i < 2; i++) {
      HIPCK(hipHostMalloc((void **)&buf[i], STAGE, hipHostMallocDefault));
      HIPCK(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    }
  }
};
Stage &stage() {
  static Stage *s = new Stage();  // kept until the process ends
  return *s;
}

void put_marker(FILE *f, size_t bytes) {
  if (bytes > 0x7fffffffUL) throw Error(1, "checkpoint record longer than 2 GiB");
  const int m = (int)bytes;
  if (fwrite(&m, 4, 1, f) != 1) throw Error(1, "checkpoint: write error");
}

// nrec records of rec bytes each, from p (host or device), consecutive in memory
void put_records(FILE *f, const void *p, size_t rec, size_t nrec) {
  if (!is_device_ptr(p)) {
    for (size_t r = 0; r < nrec; r++) {
      put_marker(f, rec);
      if (rec && fwrite((const char *)p + rec * r, 1, rec, f) != rec) throw Error(1, "checkpoint: write error");
      put_marker(f, rec);
    }
    return;
  }
  Stage &st = stage();
  hipStream_t s = lib().stream;
  const size_t cpr = rec ? (rec + STAGE - 1) / STAGE : 0;  // chunks per record
  const size_t total = cpr * nrec;
  auto chunk = [&](size_t c, size_t *off, size_t *len) {
    const size_t r = c / cpr, k = c % cpr;
    *off = rec * r + STAGE * k;
    *len = std::min(STAGE, rec - STAGE * k);
  };
  auto issue = [&](size_t c) {
    size_t off, len;
    chunk(c, &off, &len);
    HIPCK(hipMemcpyAsync(st.buf[c % 2], (const char *)p + off, len, hipMemcpyDeviceToHost, s));
    HIPCK(hipEventRecord(st.ev[c % 2], s));
  };
  if (total == 0) {
    for (size_t r = 0; r < nrec; r++) { put_marker(f, rec); put_marker(f, rec); }
    return;
  }
  issue(0);
  for (size_t c = 0; c < total; c++) {
    if (c + 1 < total) issue(c + 1);  // its buffer's previous chunk was written last iteration
    HIPCK(hipEventSynchronize(st.ev[c % 2]));
    size_t off, len;
    chunk(c, &off, &len);
    if (c % cpr == 0) put_marker(f, rec);
    if (fwrite(st.buf[c % 2], 1, len, f) != len) throw Error(1, "checkpoint: write error");
    if (c % cpr == cpr - 1) put_marker(f, rec);
  }
}

// the first `want` bytes of each of nrec records into p (host or device); a shorter record
// or a broken frame is the reference's err= branch (message err)
void get_records(FILE *f, void *p, size_t want, size_t nrec, const char *err) {
  Stage &st = stage();
  const bool dev = is_device_ptr(p);
  hipStream_t s = lib().stream;
  size_t used = 0;  // chunks issued, for the buffer rotation
  for (size_t r = 0; r < nrec; r++) {
    int m = 0, m2 = 0;
    if (fread(&m, 4, 1, f) != 1 || m < 0 || (size_t)m < want) throw Error(1, err);
    char *dst = (char *)p + want * r;
    if (!dev) {
      if (want && fread(dst, 1, want, f) != want) throw Error(1, err);
    } else {
      for (size_t off = 0; off < want; off += STAGE, used++) {
        const size_t len = std::min(STAGE, want - off);
        const int b = (int)(used % 2);
        if (used >= 2) HIPCK(hipEventSynchronize(st.ev[b]));  // its previous copy is done
        if (fread(st.buf[b], 1, len, f) != len) throw Error(1, err);
        HIPCK(hipMemcpyAsync(dst + off, st.buf[b], len, hipMemcpyHostToDevice, s));
        HIPCK(hipEventRecord(st.ev[b], s));
      }
    }
    if (fseek(f, (long)((size_t)m - want), SEEK_CUR) != 0) throw Error(1, err);
    if (fread(&m2, 4, 1, f) != 1 || m2 != m) throw Error(1, err);
  }
  if (dev) HIPCK(hipStreamSynchronize(s));
}
void skip_record(FILE *f, const char *err) { get_records(f, nullptr, 0, 1, err); }
template <class T>
T get_scalar_rec(FILE *f, const char *err) {
  T v{};
  get_records(f, &v, sizeof v, 1, err);
  return v;
}

// band-sized operations on host or device arrays
__global__ void k_real_to_complex(long n, const double *re, double2 *out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double2(re[i], 0.0);
}
__global__ void k_conj_copy(long n, const double2 *src, double2 *dst) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = make_double2(src[i].x, -src[i].y);
}
struct Bands {
  bool dev;
  long n;
  void copy(void *dst, const void *src, size_t bytes) const {
    if (dev) HIPCK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, lib().stream));
    else memmove(dst, src, bytes);
  }
  void zero(void *dst, size_t bytes) const {
    if (dev) HIPCK(hipMemsetAsync(dst, 0, bytes, lib().stream));
    else memset(dst, 0, bytes);
  }
  void complexify(const double *re, double *out) const {
    if (dev) {
      hipLaunchKernelGGL(k_real_to_complex, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, lib().stream, n,
                         re, (double2 *)out);
      HIPCK(hipGetLastError());
    } else {
      for (long i = 0; i < n; i++) { out[2 * i] = re[i]; out[2 * i + 1] = 0.0; }
    }
  }
  void conj_copy(const double *src, double *dst) const {
    if (dev) {
      hipLaunchKernelGGL(k_conj_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, lib().stream, n,
                         (const double2 *)src, (double2 *)dst);
      HIPCK(hipGetLastError());
    } else {
      for (long i = 0; i < n; i++) { dst[2 * i] = src[2 * i]; dst[2 * i + 1] = -src[2 * i + 1]; }
    }
  }
};

// a small copy between host and / or device arrays, ordered with the library stream
void get_host_or_dev(void *dst, const void *src, size_t bytes) {
  if (!bytes) return;
  if (is_device_ptr(dst) || is_device_ptr(src)) {
    HIPCK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, lib().stream));
    HIPCK(hipStreamSynchronize(lib().stream));
  } else {
    memcpy(dst, src, bytes);
  }
}

}  // namespace

extern "C" {

int kukan_write_wf(int nr1In, int nr2In, const char *chdir, int nrc, int nums, int ncol, int neigmx,
                   int numk, int ncpx, int ncpy, int ncpz, int nrxIn, int nryIn, int nrzIn, int ksym,
                   const double *svecre, const void *sveccm, const double *sval) {
  API_BEGIN
  const std::string path = ckpt_path(chdir, "wfc", nr1In < 0 ? 0 : nr1In, nr2In < 0 ? 0 : nr2In);
  File fl(path, "wb");
  if (!fl.f) throw Error(1, "cannot open wfc.* file(s)");
  const int dims[8] = {nrc, nums, ncol, neigmx, numk, ncpx, ncpy, ncpz};
  const int ranks[4] = {nrxIn < 0 ? 0 : nrxIn, nryIn < 0 ? 0 : nryIn, nrzIn < 0 ? 0 : nrzIn, 0};
  const size_t n = (size_t)ncpx * ncpy * ncpz;
  put_records(fl.f, &WFC_VERSION, 4, 1);
  put_records(fl.f, dims, sizeof dims, 1);
  put_records(fl.f, ranks, sizeof ranks, 1);
  put_records(fl.f, &ksym, 4, 1);
  put_records(fl.f, sval, 8 * (size_t)neigmx * (nums + 1 - ncol) * numk, 1);
  put_records(fl.f, nrc ? sveccm : (const void *)svecre, (nrc ? 16 : 8) * n, (size_t)neigmx * nums * numk);
  API_END
}

int kukan_read_wf(int nr1In, int nr2In, const char *chdir, int nrc, int nums, int ncol, int neigmx,
                  int numk, int ncpx, int ncpy, int ncpz, int ksym, int key_ksym_inv2, double *svecre,
                  void *sveccm, double *sval) {
  API_BEGIN
  const char *E = "error found in wfc.*";
  const std::string path = ckpt_path(chdir, "wfc", nr1In < 0 ? 0 : nr1In, nr2In < 0 ? 0 : nr2In);
  File fl(path, "rb");
  if (!fl.f) throw Error(1, "cannot open wfc.* file(s)");
  FILE *f = fl.f;
  if (get_scalar_rec<int>(f, E) != WFC_VERSION) throw Error(1, "wrong wfc version");
  int d[8];
  get_records(f, d, sizeof d, 1, E);
  skip_record(f, E);
  (void)get_scalar_rec<int>(f, E);  // ksym of the file
  const int nrc1 = d[0], nums1 = d[1], ncol1 = d[2], neigmx1 = d[3], numk1 = d[4];
  if (d[5] != ncpx || d[6] != ncpy || d[7] != ncpz || nrc1 > nrc || nums1 > nums || ncol1 > ncol ||
      (numk1 != numk && (numk1 * 2 != numk || ksym != key_ksym_inv2)))
    throw Error(1, "read_wf: unexpected array dimensions in file " + path);
  if (neigmx1 * ncol < neigmx * ncol1) throw Error(1, "read_wf: not enough eigenstates in file " + path);
  const long n = (long)ncpx * ncpy * ncpz;
  const int d2 = nums + 1 - ncol;
  const size_t el = nrc ? 16 : 8;
  char *V = nrc ? (char *)sveccm : (char *)svecre;
  auto band = [&](int l, int ns, int nk) {
    return V + el * n * ((l - 1) + (size_t)neigmx * ((ns - 1) + (size_t)nums * (nk - 1)));
  };
  if (nrc1 == nrc && nums1 == nums && ncol1 == ncol && neigmx1 == neigmx && numk1 == numk) {
    get_records(f, sval, 8 * (size_t)neigmx * d2 * numk, 1, E);
    get_records(f, V, el * n, (size_t)neigmx * nums * numk, E);  // straight into place
    sync();
    return 0;
  }
  // the mapping of read_wf (rw_wfpot8f.F90:70-178): sval on the host, bands where they live
  const bool dev = is_device_ptr(V);
  Bands B{dev, n};
  const size_t nsv = (size_t)neigmx * d2 * numk;
  std::vector<double> sv(nsv);
  get_host_or_dev(sv.data(), sval, 8 * nsv);
  auto SV = [&](int l, int ns, int nk) -> double & {
    return sv[(l - 1) + (size_t)neigmx * ((ns - 1) + (size_t)d2 * (nk - 1))];
  };
  const int d21 = nums1 + 1 - ncol1;
  std::vector<double> sval1((size_t)neigmx1 * d21 * numk1);
  auto SV1 = [&](int l, int ns, int nk) {
    return sval1[(l - 1) + (size_t)neigmx1 * ((ns - 1) + (size_t)d21 * (nk - 1))];
  };
  get_records(f, sval1.data(), 8 * sval1.size(), 1, E);
  // one file band at a time: its record into tmp (where the bands live), converted to complex
  // in tmpc when the run is complex and the file real
  double *tmp = dev ? (double *)scratch("ckpt.band", 16 * n) : nullptr;
  double *tmpc = dev ? (double *)scratch("ckpt.bandc", 16 * n) : nullptr;
  std::vector<double> htmp, htmpc;
  if (!dev) { htmp.resize(2 * n); htmpc.resize(2 * n); tmp = htmp.data(); tmpc = htmpc.data(); }
  auto read_band = [&]() -> const double * {  // the band as the run's element type
    get_records(f, tmp, (nrc1 ? 16 : 8) * n, 1, E);
    if (nrc1 < nrc) {
      B.complexify(tmp, tmpc);
      return tmpc;
    }
    return tmp;
  };
  if (ncol1 == ncol || nums1 < nums) {
    for (int nk = 1; nk <= numk1; nk++)
      for (int ns = 1; ns <= nums1; ns++)
        for (int l1 = 1; l1 <= neigmx1; l1++) {
          const int l = ncol1 == ncol ? l1 : 2 * l1 - 1;
          if (l <= neigmx && ns <= d2) SV(l, ns, nk) = SV1(l1, ns, nk);
          const double *bnd = read_band();
          if (l > neigmx) continue;
          B.copy(band(l, ns, nk), bnd, el * n);
          if (nums1 < nums) {
            if (ncol1 == ncol) {
              SV(l, nums, nk) = SV(l, 1, nk);
              B.copy(band(l, nums, nk), band(l, 1, nk), el * n);
            } else {
              B.zero(band(l, nums, nk), el * n);
              if (l + 1 <= neigmx) {
                SV(l + 1, 1, nk) = SV(l, 1, nk);
                B.copy(band(l + 1, nums, nk), band(l, 1, nk), el * n);
                B.zero(band(l + 1, 1, nk), el * n);
              }
            }
          }
          if (dev) sync();  // tmp is reused by the next band
        }
  } else {
    // a collinear two-spin file into a noncollinear run: bands merged by energy
    std::vector<int> map((size_t)nums1 * neigmx1);
    for (int nk = 1; nk <= numk1; nk++) {
      std::fill(map.begin(), map.end(), 0);
      int lspin[2] = {1, 1};
      for (int l = 1; l <= neigmx; l++) {
        int ns = 0;
        if (lspin[0] > neigmx1) ns = 2;
        if (lspin[1] > neigmx1) ns = 1;
        if (ns == 0) ns = SV1(lspin[0], 1, nk) <= SV1(lspin[1], 2, nk) ? 1 : 2;
        map[(ns - 1) + (size_t)nums1 * (lspin[ns - 1] - 1)] = l;
        lspin[ns - 1]++;
      }
      for (int ns = 1; ns <= nums1; ns++)
        for (int l1 = 1; l1 <= neigmx1; l1++) {
          const int l = map[(ns - 1) + (size_t)nums1 * (l1 - 1)];
          const double *bnd = read_band();
          if (l > 0) {
            SV(l, 1, nk) = SV1(l1, ns, nk);
            B.copy(band(l, ns, nk), bnd, el * n);
            B.zero(band(l, 3 - ns, nk), el * n);
          }
          if (dev) sync();
        }
    }
  }
  if (numk1 * 2 == numk) {
    if (ksym != key_ksym_inv2) throw Error(1, "read_wf: wrong ksym for k-point mapping");
    for (int nk = 1; nk <= numk1; nk++) {
      const int nk2 = 2 * numk1 + 1 - nk;
      for (int ns = 1; ns <= d2; ns++)
        for (int l = 1; l <= neigmx; l++) SV(l, ns, nk2) = SV(l, ns, nk);
      for (int ns = 1; ns <= nums; ns++)
        for (int l = 1; l <= neigmx; l++) {
          if (nrc == 0) B.copy(band(l, ns, nk2), band(l, ns, nk), 8 * n);
          else B.conj_copy((const double *)band(l, ns, nk), (double *)band(l, ns, nk2));
        }
    }
  }
  get_host_or_dev(sval, sv.data(), 8 * nsv);
  API_END
}

int kukan_read_rho(int nrIn, const char *chdir, int nevhist, int nspv, int ncpx, int ncpy, int ncpz,
                   int nradmx, int npoint, int num_atcell, int *lbrydn, int *na_atcell,
                   int *natcell_inf, double *rhosmt, double *rhotrur, double *rhosmtr) {
  API_BEGIN
  bool histyn = nevhist == 1 || nevhist == 3 || nevhist == -1 || nevhist == 4;
  const char *E = "error found in rho.*";
  if (histyn) {
    const std::string path = ckpt_path(chdir, nevhist == 4 ? "r_o" : "rho", nrIn < 0 ? 0 : nrIn, -1);
    File fl(path, "rb");
    histyn = fl.f != nullptr;
    if (nevhist == -1 && !histyn) throw Error(1, "read_rho: nevhist=-1, but no density file rho.*");
    if (histyn) {
      FILE *f = fl.f;
      *lbrydn = 1;
      if (get_scalar_rec<int>(f, E) != RHO_VERSION) throw Error(1, "wrong rho version");
      int d[7];
      get_records(f, d, sizeof d, 1, E);
      skip_record(f, E);
      const int nspv1 = d[0], num_atcell1 = d[6];
      std::vector<int> rec((size_t)num_atcell1 + 1);
      get_records(f, rec.data(), 4 * rec.size(), 1, E);
      *na_atcell = rec[0];
      if (nspv1 > nspv && nspv1 == 2)
        throw Error(1, "read_rho:  mapping of (mag dens) to (non-mag dens) not implemented");
      if (nspv1 > nspv && nspv1 == 4)
        throw Error(1, "read_rho:  mapping of (noco dens) to (col dens) not implemented");
      if (nspv1 > nspv || d[1] != ncpx || d[2] != ncpy || d[3] != ncpz || d[4] != nradmx || d[5] != npoint)
        throw Error(1, "read_rho: unexpected array dimensions in density file " + path);
      if (*na_atcell > num_atcell) throw Error(1, "read_rho:  num_atcell is too small");
      get_host_or_dev(natcell_inf, rec.data() + 1, 4 * (size_t)std::max(0, *na_atcell));
      const size_t n = (size_t)ncpx * ncpy * ncpz, ra = (size_t)nradmx * npoint;
      if (nspv1 == nspv) {
        get_records(f, rhosmt, 8 * n * nspv, 1, E);
        if (num_atcell1 <= num_atcell) {
          get_records(f, rhotrur, 8 * ra * nspv * num_atcell1, 1, E);
          get_records(f, rhosmtr, 8 * ra * nspv * num_atcell1, 1, E);
        } else {
          std::vector<double> t(ra * nspv * num_atcell1);
          get_records(f, t.data(), 8 * t.size(), 1, E);
          get_host_or_dev(rhotrur, t.data(), 8 * ra * nspv * (size_t)*na_atcell);
          get_records(f, t.data(), 8 * t.size(), 1, E);
          get_host_or_dev(rhosmtr, t.data(), 8 * ra * nspv * (size_t)*na_atcell);
        }
      } else {
        // spin expansion on the host (densities are a few grid vectors)
        std::vector<double> s1(n * nspv1), t1(ra * nspv1 * num_atcell1), u1(ra * nspv1 * num_atcell1);
        get_records(f, s1.data(), 8 * s1.size(), 1, E);
        get_records(f, t1.data(), 8 * t1.size(), 1, E);
        get_records(f, u1.data(), 8 * u1.size(), 1, E);
        std::vector<double> s(n * nspv), t(ra * nspv * num_atcell), u(ra * nspv * num_atcell);
        get_host_or_dev(t.data(), rhotrur, 8 * t.size());
        get_host_or_dev(u.data(), rhosmtr, 8 * u.size());
        for (int ns = 1; ns <= nspv; ns++) {
          const int ns1 = std::min(ns, nspv1);
          double fctr = 1.0;
          if (nspv1 == 1) fctr = 0.5;
          if (ns > 2) fctr = 0.0;
          for (size_t i = 0; i < n; i++) s[i + n * (ns - 1)] = s1[i + n * (ns1 - 1)] * fctr;
          for (int na = 0; na < *na_atcell; na++)
            for (size_t i = 0; i < ra; i++) {
              t[i + ra * ((ns - 1) + (size_t)nspv * na)] = t1[i + ra * ((ns1 - 1) + (size_t)nspv1 * na)] * fctr;
              u[i + ra * ((ns - 1) + (size_t)nspv * na)] = u1[i + ra * ((ns1 - 1) + (size_t)nspv1 * na)] * fctr;
            }
        }
        get_host_or_dev(rhosmt, s.data(), 8 * s.size());
        get_host_or_dev(rhotrur, t.data(), 8 * t.size());
        get_host_or_dev(rhosmtr, u.data(), 8 * u.size());
      }
    }
  }
  if (!histyn) *lbrydn = 0;
  API_END
}

int kukan_read_vht(int nrIn, const char *chdir, int nevhist, int ncpx, int ncpy, int ncpz, int nmesh,
                   double *vh_dense) {
  API_BEGIN
  const size_t nd = (size_t)ncpx * nmesh * ncpy * nmesh * ncpz * nmesh;
  bool histyn = nevhist == 2 || nevhist == 3 || nevhist == -1;
  const char *E = "error found in vht.*";
  if (histyn) {
    const std::string path = ckpt_path(chdir, "vht", nrIn < 0 ? 0 : nrIn, -1);
    File fl(path, "rb");
    histyn = fl.f != nullptr;
    if (histyn) {
      if (get_scalar_rec<int>(fl.f, E) != VHT_VERSION) throw Error(1, "wrong vht version");
      int d[4];
      get_records(fl.f, d, sizeof d, 1, E);
      skip_record(fl.f, E);
      if (d[0] != ncpx || d[1] != ncpy || d[2] != ncpz || d[3] != nmesh)
        throw Error(1, "read_vht: unexpected array dimensions in potential file " + path);
      get_records(fl.f, vh_dense, 8 * nd, 1, E);
    }
  }
  if (!histyn) {
    if (is_device_ptr(vh_dense)) launch_fill((long)nd, 0.0, vh_dense, lib().stream);
    else memset(vh_dense, 0, 8 * nd);
  }
  API_END
}

int kukan_write_rhovht(int nrIn, const char *chdir, int nrho, int nele, int npre, int nevhist, int nspv,
                       int ncpx, int ncpy, int ncpz, int nmesh, int nradmx, int npoint, int num_atcell,
                       int nrxIn, int nryIn, int nrzIn, int natom, int key_natpri_in, const int *natpri,
                       const double *rhosmt_i, const double *rhotrur_i, const double *rhosmtr_i,
                       const double *rhosmt_o, const double *rhotrur_o, const double *rhosmtr_o,
                       const double *vh_dense) {
  API_BEGIN
  const int nr = nrIn < 0 ? 0 : nrIn;
  const int rk[3] = {nrxIn < 0 ? 0 : nrxIn, nryIn < 0 ? 0 : nryIn, nrzIn < 0 ? 0 : nrzIn};
  const size_t n = (size_t)ncpx * ncpy * ncpz, ra = (size_t)nradmx * npoint;
  if (nevhist == 1 || nevhist == 3) {
    const std::vector<int> np_ = host_ints(natpri, natom);
    for (int irho = 1; irho <= nrho; irho++) {
      const std::string path = ckpt_path(chdir, irho == 1 ? "rho" : "r_o", nr, -1);
      File fl(path, "wb");
      if (!fl.f) throw Error(1, "cannot open rho.* file(s)");
      if (npre == 2) {
        fclose(fl.f);
        fl.f = nullptr;
        remove(path.c_str());
        continue;
      }
      std::vector<int> rec((size_t)num_atcell + 1, 0);
      int na_atcell = 0;
      for (int na = 1; na <= natom; na++)
        if (np_[na - 1] == key_natpri_in) {
          na_atcell++;
          if (na_atcell > num_atcell) throw Error(1, "write_rhovht: na_atcell > num_atcell");
          rec[na_atcell] = na;
        }
      rec[0] = na_atcell;
      const int d[7] = {nspv, ncpx, ncpy, ncpz, nradmx, npoint, na_atcell};  // num_atcell1 (l_samedim false)
      put_records(fl.f, &RHO_VERSION, 4, 1);
      put_records(fl.f, d, sizeof d, 1);
      put_records(fl.f, rk, sizeof rk, 1);
      put_records(fl.f, rec.data(), 4 * rec.size(), 1);
      put_records(fl.f, irho == 1 ? rhosmt_i : rhosmt_o, 8 * n * nspv, 1);
      put_records(fl.f, irho == 1 ? rhotrur_i : rhotrur_o, 8 * ra * nspv * na_atcell, 1);
      put_records(fl.f, irho == 1 ? rhosmtr_i : rhosmtr_o, 8 * ra * nspv * na_atcell, 1);
    }
  }
  if (nevhist == 2 || nevhist == 3) {
    const std::string path = ckpt_path(chdir, "vht", nr, -1);
    File fl(path, "wb");
    if (!fl.f) throw Error(1, "cannot open vht.* file(s)");
    if (npre == 2) {
      fclose(fl.f);
      fl.f = nullptr;
      remove(path.c_str());
    } else {
      const int d[4] = {ncpx, ncpy, ncpz, nmesh};
      put_records(fl.f, &VHT_VERSION, 4, 1);
      put_records(fl.f, d, sizeof d, 1);
      put_records(fl.f, rk, sizeof rk, 1);
      put_records(fl.f, vh_dense, 8 * n * (size_t)nmesh * nmesh * nmesh, 1);
    }
  }
  if (nele == 1) {
    File fl(ckpt_path(chdir, "ele", nr, -1), "wb");
    if (!fl.f) throw Error(1, "cannot open ele.* file(s)");
    put_records(fl.f, rhosmt_o, 8 * n * nspv, 1);
  }
  API_END
}

int kukan_filenumbers_rhovht(const char *chdir, int *nrrho, int *nrpot) {
  API_BEGIN
  const char *names[2] = {"rho", "vht"};
  int *outs[2] = {nrrho, nrpot};
  for (int k = 0; k < 2; k++) {
    int nr = 0;
    struct stat sb;
    while (stat(ckpt_path(chdir, names[k], nr, -1).c_str(), &sb) == 0) nr++;
    *outs[k] = nr;
  }
  API_END
}

}  // extern "C"
