// plot3d.cpp -- PLOT3D whole-format files and actuator control / gradient files
// (host only, no device code).
//
// PLOT3D (src/PLOT3DFormat.c:88-294, src/PLOT3DHelperImpl.f90:554-1286): Fortran
// unformatted sequential records, each framed by a leading and trailing byte count
// (4-byte markers, or 8-byte ones in the reference's USE_EXTENDED_PLOT3D build,
// include/config.h.in:34-40), in the byte order of the writer:
//   record 1  number of grids (int32)
//   record 2  per grid nx, ny, nz (int32), plus the number of scalars for function files
//   then per grid
//     grid      x, y, z (N doubles each) and iblank (N int32) -- iblank may be absent
//     solution  a 4-double auxiliary record (timestep, -, -, time), then 5 components
//               (rho, rho u, rho v, rho w, rho E); a 2-D state fills slots 1-3 and 5
//     function  nScalars components
// The byte order and the marker width are detected from the first record: its marker
// reads 4; with 4-byte markers the next 4 bytes are the (positive) grid count, with 8-byte
// markers they are the zero high half of the marker.  Writers use 4-byte markers unless
// mgd_plot3d_set_record_markers(8) was called.
//
// Control forcing and gradient files (src/ActuatorPatchImpl.f90:226-458): raw native
// doubles, one record of nPatchPoints values per RK stage, in the order the adjoint
// march writes them (latest stage first).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mgd_internal.h"

namespace {

int io_error(const std::string& m, int rc = MGD_ERR_ARG) {
  mgd::set_error(m);
  return rc;
}

uint32_t swap32(uint32_t x) {
  return ((x >> 24) & 0xffu) | ((x >> 8) & 0xff00u) | ((x << 8) & 0xff0000u) | ((x << 24) & 0xff000000u);
}

uint64_t swap64(uint64_t x) {
  uint64_t r = 0;
  for (int i = 0; i < 8; ++i) r = (r << 8) | ((x >> (8 * i)) & 0xffu);
  return r;
}

int g_marker_bytes = 4;   // record markers the writers use

struct Reader {
  std::vector<unsigned char> buf;
  size_t pos = 0;
  bool native = true;
  bool wide = false;         // 8-byte record markers
  bool load(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize(n > 0 ? (size_t)n : 0);
    const size_t got = buf.empty() ? 0 : std::fread(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    return got == buf.size();
  }
  bool i32(int32_t* v) {
    if (pos + 4 > buf.size()) return false;
    uint32_t u;
    std::memcpy(&u, &buf[pos], 4);
    pos += 4;
    if (!native) u = swap32(u);
    std::memcpy(v, &u, 4);
    return true;
  }
  bool f64(double* v, size_t n) {
    if (pos + 8 * n > buf.size()) return false;
    for (size_t i = 0; i < n; ++i) {
      uint64_t u;
      std::memcpy(&u, &buf[pos + 8 * i], 8);
      if (!native) u = swap64(u);
      std::memcpy(&v[i], &u, 8);
    }
    pos += 8 * n;
    return true;
  }
  bool i64(int64_t* v) {
    if (pos + 8 > buf.size()) return false;
    uint64_t u;
    std::memcpy(&u, &buf[pos], 8);
    pos += 8;
    if (!native) u = swap64(u);
    std::memcpy(v, &u, 8);
    return true;
  }
  bool marker(int64_t* m) {
    if (wide) return i64(m);
    int32_t v;
    if (!i32(&v)) return false;
    *m = v;
    return true;
  }
  // a whole record: returns its byte count and leaves pos at its first byte
  bool open_record(int64_t* bytes) {
    int64_t m;
    if (!marker(&m)) return false;
    *bytes = m;
    return m >= 0 && pos + (size_t)m + (wide ? 8 : 4) <= buf.size();
  }
  bool close_record(int64_t bytes) {
    int64_t m;
    return marker(&m) && m == bytes;
  }
};

struct Layout {
  int n_grids = 0, file_type = 0, n_scalars = 0, has_iblank = 0;
  std::vector<int32_t> sizes;   // 3 per grid
  std::vector<size_t> data_pos; // byte offset of each grid's first data record (its marker)
};

// Record structure of a PLOT3D file (detect_format, src/PLOT3DFormat.c:88-205).
int parse(Reader& R, int include_function, Layout* L) {
  if (R.buf.size() < 12) return io_error("empty or truncated PLOT3D file", MGD_ERR_STATE);
  int32_t first, second;
  std::memcpy(&first, R.buf.data(), 4);
  std::memcpy(&second, R.buf.data() + 4, 4);
  int64_t first8;
  std::memcpy(&first8, R.buf.data(), 8);
  if ((first == 4 || (int32_t)swap32((uint32_t)first) == 4) && second != 0) {
    R.wide = false;
    R.native = first == 4;
  } else if (first8 == 4 || (int64_t)swap64((uint64_t)first8) == 4) {
    R.wide = true;
    R.native = first8 == 4;
  } else {
    return io_error("not a PLOT3D whole-format file (first record marker is not 4)", MGD_ERR_STATE);
  }
  // every record: matching leading and trailing markers, last one followed by EOF
  R.pos = 0;
  while (R.pos < R.buf.size()) {
    int64_t b;
    if (!R.open_record(&b)) return io_error("PLOT3D record marker runs past the end of file", MGD_ERR_STATE);
    R.pos += (size_t)b;
    if (!R.close_record(b)) return io_error("PLOT3D leading and trailing record sizes differ", MGD_ERR_STATE);
  }
  R.pos = 0;
  int64_t b;
  int32_t ng;
  if (!R.open_record(&b) || !R.i32(&ng) || !R.close_record(b) || ng <= 0)
    return io_error("bad PLOT3D grid count", MGD_ERR_STATE);
  L->n_grids = ng;
  if (!R.open_record(&b)) return io_error("missing PLOT3D size record", MGD_ERR_STATE);
  if (b != 12LL * ng && !(include_function && b == 16LL * ng))
    return io_error("bad PLOT3D size record", MGD_ERR_STATE);
  L->file_type = b == 16LL * ng ? 2 : 0;
  const int per = L->file_type == 2 ? 4 : 3;
  L->sizes.assign(3 * (size_t)ng, 0);
  for (int g = 0; g < ng; ++g)
    for (int c = 0; c < per; ++c) {
      int32_t v;
      if (!R.i32(&v) || v <= 0) return io_error("bad PLOT3D grid size", MGD_ERR_STATE);
      if (c < 3) L->sizes[3 * g + c] = v;
      else if (g == 0) L->n_scalars = v;
      else if (v != L->n_scalars) return io_error("function components differ between grids", MGD_ERR_STATE);
    }
  if (!R.close_record(b)) return io_error("bad PLOT3D size record", MGD_ERR_STATE);
  L->data_pos.assign((size_t)ng, 0);
  for (int g = 0; g < ng; ++g) {
    const int64_t N = (int64_t)L->sizes[3 * g] * L->sizes[3 * g + 1] * L->sizes[3 * g + 2];
    L->data_pos[g] = R.pos;
    if (!R.open_record(&b)) return io_error("truncated PLOT3D file", MGD_ERR_STATE);
    if (L->file_type == 2) {
      if (b != 8LL * L->n_scalars * N) return io_error("bad PLOT3D function record", MGD_ERR_STATE);
    } else if (b == 32) {   // auxiliary record: solution file
      L->file_type = 1;
      R.pos += 32;
      if (!R.close_record(32) || !R.open_record(&b) || b != 40LL * N)
        return io_error("bad PLOT3D solution record", MGD_ERR_STATE);
    } else {
      if (b != 24LL * N && b != 28LL * N) return io_error("bad PLOT3D grid record", MGD_ERR_STATE);
      L->has_iblank = b == 28LL * N;
    }
    R.pos += (size_t)b;
    if (!R.close_record(b)) return io_error("bad PLOT3D record trailer", MGD_ERR_STATE);
  }
  return MGD_OK;
}

int open_block(const char* path, int block, int want_type, Reader& R, Layout& L, int64_t* N) {
  if (!path) return io_error("null path");
  if (!R.load(path)) return io_error(std::string("cannot read ") + path, MGD_ERR_STATE);
  if (int rc = parse(R, want_type == 2, &L)) return rc;
  if (L.file_type != want_type) return io_error("PLOT3D file is of another type", MGD_ERR_ARG);
  if (block < 0 || block >= L.n_grids) return io_error("PLOT3D block out of range", MGD_ERR_ARG);
  *N = (int64_t)L.sizes[3 * block] * L.sizes[3 * block + 1] * L.sizes[3 * block + 2];
  R.pos = L.data_pos[(size_t)block];
  return MGD_OK;
}

struct Writer {
  FILE* f = nullptr;
  bool ok = true;
  explicit Writer(const char* path) { f = std::fopen(path, "wb"); ok = f != nullptr; }
  ~Writer() { if (f) std::fclose(f); }
  void raw(const void* p, size_t n) { if (ok) ok = std::fwrite(p, 1, n, f) == n; }
  void marker(int64_t b) {
    if (g_marker_bytes == 8) {
      raw(&b, 8);
    } else {
      const int32_t m = (int32_t)b;
      raw(&m, 4);
    }
  }
};

void write_header(Writer& W, const int* n, int nscalars) {
  W.marker(4);
  const int32_t one = 1;
  W.raw(&one, 4);
  W.marker(4);
  const int64_t b = nscalars > 0 ? 16 : 12;
  W.marker(b);
  for (int i = 0; i < 3; ++i) { const int32_t v = n[i]; W.raw(&v, 4); }
  if (nscalars > 0) { const int32_t v = nscalars; W.raw(&v, 4); }
  W.marker(b);
}

bool sizes_ok(const int* n) { return n && n[0] > 0 && n[1] > 0 && n[2] > 0; }

}  // namespace

extern "C" {

int mgd_plot3d_set_record_markers(int bytes) {
  if (bytes != 4 && bytes != 8) return io_error("PLOT3D record markers are 4 or 8 bytes");
  g_marker_bytes = bytes;
  return MGD_OK;
}

int mgd_plot3d_detect(const char* path, int include_function, int* n_grids, int* file_type,
                      int* n_scalars, int* has_iblank, int* native_endian, int* sizes, int max_grids) {
  Reader R;
  Layout L;
  if (!path) return io_error("null path");
  if (!R.load(path)) return io_error(std::string("cannot read ") + path, MGD_ERR_STATE);
  if (int rc = parse(R, include_function, &L)) return rc;
  if (n_grids) *n_grids = L.n_grids;
  if (file_type) *file_type = L.file_type;
  if (n_scalars) *n_scalars = L.n_scalars;
  if (has_iblank) *has_iblank = L.has_iblank;
  if (native_endian) *native_endian = R.native ? 1 : 0;
  if (sizes)
    for (int g = 0; g < L.n_grids && g < max_grids; ++g)
      for (int c = 0; c < 3; ++c) sizes[3 * g + c] = L.sizes[3 * g + c];
  return MGD_OK;
}

int mgd_plot3d_read_grid(const char* path, int block, double* coords, int* iblank) {
  Reader R;
  Layout L;
  int64_t N;
  if (int rc = open_block(path, block, 0, R, L, &N)) return rc;
  int64_t b;
  R.open_record(&b);
  if (!coords || !R.f64(coords, (size_t)(3 * N))) return io_error("bad grid data", MGD_ERR_STATE);
  if (iblank) {
    if (L.has_iblank) {
      for (int64_t i = 0; i < N; ++i) {
        int32_t v;
        R.i32(&v);
        iblank[i] = v;
      }
    } else {
      for (int64_t i = 0; i < N; ++i) iblank[i] = 1;
    }
  }
  return MGD_OK;
}

int mgd_plot3d_read_solution(const char* path, int block, int nv, double* state, double* aux) {
  if (nv < 3 || nv > 5) return io_error("a PLOT3D solution holds 3 to 5 state components");
  Reader R;
  Layout L;
  int64_t N;
  if (int rc = open_block(path, block, 1, R, L, &N)) return rc;
  int64_t b;
  double a[4];
  R.open_record(&b);
  R.f64(a, 4);
  R.close_record(b);
  if (aux) std::memcpy(aux, a, sizeof a);
  R.open_record(&b);
  const size_t base = R.pos;
  if (!state) return io_error("null state");
  // components 1 .. nv-1 in slots 1 .. nv-1, the energy in slot 5
  for (int l = 0; l < nv; ++l) {
    const int slot = l < nv - 1 ? l : 4;
    R.pos = base + (size_t)slot * 8 * (size_t)N;
    if (!R.f64(state + (size_t)l * N, (size_t)N)) return io_error("bad solution data", MGD_ERR_STATE);
  }
  return MGD_OK;
}

int mgd_plot3d_read_function(const char* path, int block, int n_scalars, double* data) {
  Reader R;
  Layout L;
  int64_t N;
  if (int rc = open_block(path, block, 2, R, L, &N)) return rc;
  if (n_scalars != L.n_scalars) return io_error("function file has another number of components");
  int64_t b;
  R.open_record(&b);
  if (!data || !R.f64(data, (size_t)n_scalars * N)) return io_error("bad function data", MGD_ERR_STATE);
  return MGD_OK;
}

int mgd_plot3d_write_grid(const char* path, int nd, const int* n, const double* coords,
                          const int* iblank) {
  if (!path || !coords || nd < 1 || nd > 3 || !sizes_ok(n)) return io_error("bad argument");
  const int64_t N = (int64_t)n[0] * n[1] * n[2];
  Writer W(path);
  write_header(W, n, 0);
  W.marker(28 * N);
  W.raw(coords, 8 * (size_t)(nd * N));
  const std::vector<double> zero((size_t)N, 0.0);
  for (int d = nd; d < 3; ++d) W.raw(zero.data(), 8 * (size_t)N);
  std::vector<int32_t> ib((size_t)N, 1);
  if (iblank)
    for (int64_t i = 0; i < N; ++i) ib[i] = iblank[i];
  W.raw(ib.data(), 4 * (size_t)N);
  W.marker(28 * N);
  if (!W.ok) return io_error(std::string("cannot write ") + path, MGD_ERR_STATE);
  return MGD_OK;
}

int mgd_plot3d_write_solution(const char* path, const int* n, int nv, const double* state,
                              const double* aux) {
  if (!path || !state || nv < 3 || nv > 5 || !sizes_ok(n)) return io_error("bad argument");
  const int64_t N = (int64_t)n[0] * n[1] * n[2];
  Writer W(path);
  write_header(W, n, 0);
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  if (aux) std::memcpy(a, aux, sizeof a);
  W.marker(32);
  W.raw(a, 32);
  W.marker(32);
  W.marker(40 * N);
  const std::vector<double> zero((size_t)N, 0.0);
  for (int slot = 0; slot < 5; ++slot) {
    const double* src = slot < nv - 1 ? state + (size_t)slot * N
                      : slot == 4 ? state + (size_t)(nv - 1) * N : zero.data();
    W.raw(src, 8 * (size_t)N);
  }
  W.marker(40 * N);
  if (!W.ok) return io_error(std::string("cannot write ") + path, MGD_ERR_STATE);
  return MGD_OK;
}

int mgd_plot3d_write_function(const char* path, const int* n, int n_scalars, const double* data) {
  if (!path || !data || n_scalars <= 0 || !sizes_ok(n)) return io_error("bad argument");
  const int64_t N = (int64_t)n[0] * n[1] * n[2];
  Writer W(path);
  write_header(W, n, n_scalars);
  W.marker(8LL * n_scalars * N);
  W.raw(data, 8 * (size_t)n_scalars * (size_t)N);
  W.marker(8LL * n_scalars * N);
  if (!W.ok) return io_error(std::string("cannot write ") + path, MGD_ERR_STATE);
  return MGD_OK;
}

int mgd_control_write(const char* path, int nstages, int npoints, const double* forward_order) {
  if (!path || !forward_order || nstages < 0 || npoints <= 0) return io_error("bad argument");
  Writer W(path);
  for (int s = nstages - 1; s >= 0; --s) W.raw(forward_order + (size_t)s * npoints, 8 * (size_t)npoints);
  if (!W.ok) return io_error(std::string("cannot write ") + path, MGD_ERR_STATE);
  return MGD_OK;
}

int mgd_control_read(const char* path, int nstages, int npoints, double* forward_order) {
  if (!path || !forward_order || nstages < 0 || npoints <= 0) return io_error("bad argument");
  Reader R;
  if (!R.load(path)) return io_error(std::string("cannot read ") + path, MGD_ERR_STATE);
  if (R.buf.size() != 8 * (size_t)nstages * npoints)
    return io_error("control file size does not match stages x patch points", MGD_ERR_ARG);
  for (int s = 0; s < nstages; ++s)
    std::memcpy(forward_order + (size_t)s * npoints,
                &R.buf[8 * (size_t)(nstages - 1 - s) * npoints], 8 * (size_t)npoints);
  return MGD_OK;
}

// ZAXPY (bin/ZAXPY.f90, src/ControlSpaceAdvancerImpl.f90:3-191): z = a x + y over whole
// control files (y may be NULL: z = a x).
int mgd_control_zaxpy(const char* z, double a, const char* x, const char* y) {
  if (!z || !x) return io_error("bad argument");
  Reader X, Y;
  if (!X.load(x)) return io_error(std::string("cannot read ") + x, MGD_ERR_STATE);
  if (y && !Y.load(y)) return io_error(std::string("cannot read ") + y, MGD_ERR_STATE);
  if (X.buf.size() % 8 != 0 || (y && Y.buf.size() != X.buf.size()))
    return io_error("control files of different sizes", MGD_ERR_ARG);
  const size_t n = X.buf.size() / 8;
  std::vector<double> xv(n), yv(n, 0.0), zv(n);
  std::memcpy(xv.data(), X.buf.data(), 8 * n);
  if (y) std::memcpy(yv.data(), Y.buf.data(), 8 * n);
  for (size_t i = 0; i < n; ++i) zv[i] = y ? a * xv[i] + yv[i] : a * xv[i];
  Writer W(z);
  W.raw(zv.data(), 8 * n);
  if (!W.ok) return io_error(std::string("cannot write ") + z, MGD_ERR_STATE);
  return MGD_OK;
}

}  // extern "C"
