// On-disk formats of the hot path (src/file-io.c): the HDF5 fluid output of
// OutputFluidHDF5, the HDF5 density input of LoadFluidQuantities and the PETSc binary
// particle dump of OutputSwarmBinary.  Host code only.
//
// HDF5 is resolved at run time from the shared library (ECSPRT_HDF5_LIBRARY alone when
// set, else libhdf5.so from the loader's path, else the conda copy under /opt/conda/lib), so the
// library neither links nor needs it; every entry point returns ECSPRT_ERR_STATE with a
// message when no HDF5 1.10 or later is found.  The files follow what PETSc 3.20's
// viewers write for these objects (the reference's setup.sh pins that release):
//  - HDF5: one dataset per named vector in the root group, dimensions [Nz][Ny][Nx]
//    (three-dimensional DMDA) or [Ny][Nx] (two-dimensional), native doubles, with the
//    boolean attribute "timestepping" = 0 that VecView_MPI_HDF5_DA attaches.
//  - binary: big-endian int32 VEC_FILE_CLASSID (1211214), int32 length, the values as
//    big-endian doubles; the block size goes to "<file>.info" as
//    "-vecload_block_size <bs>".
#include <dlfcn.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace ecsprt {
namespace {

// ---- HDF5 C API, resolved at run time (1.10 and later: hid_t is 64-bit) -------------
typedef int64_t hid_t;
typedef int herr_t;
typedef int htri_t;
typedef unsigned long long hsize_t;
constexpr unsigned H5F_ACC_RDONLY = 0x0000u, H5F_ACC_TRUNC = 0x0002u;
constexpr hid_t H5P_DEFAULT = 0, H5E_DEFAULT = 0, H5S_ALL = 0;
constexpr int H5S_SCALAR = 0;

struct Hdf5 {
  bool tried = false, ok = false;
  std::string why;
  herr_t (*open)();
  herr_t (*get_libversion)(unsigned *, unsigned *, unsigned *);
  herr_t (*eset_auto2)(hid_t, void *, void *);
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*screate)(int);
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
  herr_t (*sclose)(hid_t);
  int (*sget_ndims)(hid_t);
  int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
  hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen2)(hid_t, const char *, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  hid_t (*dget_space)(hid_t);
  herr_t (*dclose)(hid_t);
  htri_t (*lexists)(hid_t, const char *, hid_t);
  hid_t (*acreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
  herr_t (*awrite)(hid_t, hid_t, const void *);
  herr_t (*aclose)(hid_t);
  hid_t t_double = -1, t_bool = -1;
};

std::mutex g_h5_mu;  // HDF5 builds are not thread-safe by default: one call at a time
Hdf5 g_h5;

template <typename F>
bool sym(void *lib, const char *name, F &f, std::string &why) {
  f = reinterpret_cast<F>(dlsym(lib, name));
  if (!f) why = std::string("HDF5 library lacks ") + name;
  return f != nullptr;
}

// Loads and checks the library once; the caller holds g_h5_mu.
Hdf5 &hdf5() {
  Hdf5 &h = g_h5;
  if (h.tried) return h;
  h.tried = true;
  // ECSPRT_HDF5_LIBRARY, when set, is the only candidate.
  std::vector<std::string> names;
  if (const char *e = std::getenv("ECSPRT_HDF5_LIBRARY")) names.push_back(e);
  else
    for (const char *n : {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so"})
      names.push_back(n);
  void *lib = nullptr;
  for (const std::string &n : names)
    if ((lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
  if (!lib) {
    h.why = "no HDF5 shared library found (ECSPRT_HDF5_LIBRARY names libhdf5.so of release 1.10 or later)";
    return h;
  }
  std::string &w = h.why;
  if (!(sym(lib, "H5open", h.open, w) && sym(lib, "H5get_libversion", h.get_libversion, w) &&
        sym(lib, "H5Eset_auto2", h.eset_auto2, w) && sym(lib, "H5Fcreate", h.fcreate, w) &&
        sym(lib, "H5Fopen", h.fopen, w) && sym(lib, "H5Fclose", h.fclose, w) && sym(lib, "H5Screate", h.screate, w) &&
        sym(lib, "H5Screate_simple", h.screate_simple, w) && sym(lib, "H5Sclose", h.sclose, w) &&
        sym(lib, "H5Sget_simple_extent_ndims", h.sget_ndims, w) &&
        sym(lib, "H5Sget_simple_extent_dims", h.sget_dims, w) && sym(lib, "H5Dcreate2", h.dcreate2, w) &&
        sym(lib, "H5Dopen2", h.dopen2, w) && sym(lib, "H5Dwrite", h.dwrite, w) && sym(lib, "H5Dread", h.dread, w) &&
        sym(lib, "H5Dget_space", h.dget_space, w) && sym(lib, "H5Dclose", h.dclose, w) &&
        sym(lib, "H5Lexists", h.lexists, w) && sym(lib, "H5Acreate2", h.acreate2, w) &&
        sym(lib, "H5Awrite", h.awrite, w) && sym(lib, "H5Aclose", h.aclose, w)))
    return h;
  unsigned maj = 0, mnr = 0, rel = 0;
  if (h.open() < 0 || h.get_libversion(&maj, &mnr, &rel) < 0 || maj != 1 || mnr < 10) {
    h.why = "HDF5 " + std::to_string(maj) + "." + std::to_string(mnr) + " found; 1.10 or later is required";
    return h;
  }
  // The native type ids are library globals, valid once H5open has run.
  const hid_t *td = static_cast<const hid_t *>(dlsym(lib, "H5T_NATIVE_DOUBLE_g"));
  const hid_t *tb = static_cast<const hid_t *>(dlsym(lib, "H5T_NATIVE_HBOOL_g"));
  if (!td || !tb) {
    h.why = "HDF5 library lacks the native type ids";
    return h;
  }
  h.t_double = *td;
  h.t_bool = *tb;
  h.eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // failures are reported through return codes
  h.ok = true;
  return h;
}

Hdf5 &need_hdf5() {
  Hdf5 &h = hdf5();
  if (!h.ok) throw Error{ECSPRT_ERR_STATE, h.why};
  return h;
}

// Closes an HDF5 handle on scope exit.
struct H5Handle {
  hid_t id;
  herr_t (*close)(hid_t);
  ~H5Handle() {
    if (id >= 0) close(id);
  }
};

// Writes named datasets of one shape into a new file (truncating an existing one).
void h5_write(const std::string &path, int nd, const hsize_t *dims, const std::vector<std::string> &names,
              const std::vector<const double *> &data) {
  std::lock_guard<std::mutex> lock(g_h5_mu);
  Hdf5 &h = need_hdf5();
  H5Handle f{h.fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), h.fclose};
  if (f.id < 0) throw Error{ECSPRT_ERR_ARG, "cannot create HDF5 file " + path};
  H5Handle space{h.screate_simple(nd, dims, nullptr), h.sclose};
  H5Handle scalar{h.screate(H5S_SCALAR), h.sclose};
  if (space.id < 0 || scalar.id < 0) throw Error{ECSPRT_ERR_ARG, "HDF5 dataspace creation failed"};
  const uint64_t timestepping = 0;  // false, whatever the size of hbool_t
  for (size_t q = 0; q < names.size(); ++q) {
    H5Handle d{h.dcreate2(f.id, names[q].c_str(), h.t_double, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
               h.dclose};
    if (d.id < 0 || h.dwrite(d.id, h.t_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, data[q]) < 0)
      throw Error{ECSPRT_ERR_ARG, "cannot write dataset \"" + names[q] + "\" to " + path};
    H5Handle a{h.acreate2(d.id, "timestepping", h.t_bool, scalar.id, H5P_DEFAULT, H5P_DEFAULT), h.aclose};
    if (a.id < 0 || h.awrite(a.id, h.t_bool, &timestepping) < 0)
      throw Error{ECSPRT_ERR_ARG, "cannot write the attributes of \"" + names[q] + "\" in " + path};
  }
}

// Reads dataset `name`: its shape into dims (at most 4 dimensions) and, when out is
// given, its values (at most cap of them) converted to double.
int h5_read(const std::string &path, const std::string &name, hsize_t *dims, double *out, long cap) {
  std::lock_guard<std::mutex> lock(g_h5_mu);
  Hdf5 &h = need_hdf5();
  H5Handle f{h.fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), h.fclose};
  if (f.id < 0) throw Error{ECSPRT_ERR_ARG, "cannot open HDF5 file " + path};
  if (h.lexists(f.id, name.c_str(), H5P_DEFAULT) <= 0)
    throw Error{ECSPRT_ERR_ARG, "no dataset \"" + name + "\" in " + path};
  H5Handle d{h.dopen2(f.id, name.c_str(), H5P_DEFAULT), h.dclose};
  if (d.id < 0) throw Error{ECSPRT_ERR_ARG, "cannot open dataset \"" + name + "\" in " + path};
  H5Handle s{h.dget_space(d.id), h.sclose};
  const int nd = s.id >= 0 ? h.sget_ndims(s.id) : -1;
  if (nd < 0 || nd > 4) throw Error{ECSPRT_ERR_ARG, "dataset \"" + name + "\" has an unsupported rank"};
  h.sget_dims(s.id, dims, nullptr);
  if (out) {
    long n = 1;
    for (int q = 0; q < nd; ++q) n *= (long)dims[q];
    if (n > cap) throw Error{ECSPRT_ERR_ARG, "dataset \"" + name + "\" holds more values than the buffer"};
    if (h.dread(d.id, h.t_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
      throw Error{ECSPRT_ERR_ARG, "cannot read dataset \"" + name + "\" from " + path};
  }
  return nd;
}

// ---- PETSc binary vectors ----------------------------------------------------------
constexpr int32_t VEC_FILE_CLASSID = 1211214;

uint32_t be32(uint32_t v) { return __builtin_bswap32(v); }
uint64_t be64(uint64_t v) { return __builtin_bswap64(v); }

void vec_write(const std::string &path, long n, int bs, const double *v) {
  if (n < 0 || n > INT32_MAX) throw Error{ECSPRT_ERR_ARG, "vector length does not fit a 32-bit PETSc index"};
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw Error{ECSPRT_ERR_ARG, "cannot create " + path};
  const uint32_t head[2] = {be32((uint32_t)VEC_FILE_CLASSID), be32((uint32_t)n)};
  bool ok = std::fwrite(head, 4, 2, f) == 2;
  std::vector<uint64_t> buf(1 << 16);
  for (long i = 0; ok && i < n; i += (long)buf.size()) {
    const long m = std::min<long>((long)buf.size(), n - i);
    for (long q = 0; q < m; ++q) {
      uint64_t u;
      std::memcpy(&u, v + i + q, 8);
      buf[q] = be64(u);
    }
    ok = std::fwrite(buf.data(), 8, m, f) == (size_t)m;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) throw Error{ECSPRT_ERR_ARG, "write to " + path + " failed"};
  if (bs > 1) {
    FILE *g = std::fopen((path + ".info").c_str(), "w");
    if (!g || std::fprintf(g, "-vecload_block_size %d\n", bs) < 0 || std::fclose(g) != 0)
      throw Error{ECSPRT_ERR_ARG, "cannot write " + path + ".info"};
  }
}

long vec_read(const std::string &path, double *v, long cap, int *bs) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) throw Error{ECSPRT_ERR_ARG, "cannot open " + path};
  uint32_t head[2];
  if (std::fread(head, 4, 2, f) != 2 || (int32_t)be32(head[0]) != VEC_FILE_CLASSID) {
    std::fclose(f);
    throw Error{ECSPRT_ERR_ARG, path + " is not a PETSc binary vector"};
  }
  const long n = (long)(int32_t)be32(head[1]);
  if (n < 0 || (v && n > cap)) {
    std::fclose(f);
    throw Error{ECSPRT_ERR_ARG, path + ": vector length " + std::to_string(n) + " exceeds the buffer"};
  }
  bool ok = true;
  if (v) {
    ok = std::fread(v, 8, n, f) == (size_t)n;
    for (long q = 0; ok && q < n; ++q) {
      uint64_t u;
      std::memcpy(&u, v + q, 8);
      u = be64(u);
      std::memcpy(v + q, &u, 8);
    }
  }
  std::fclose(f);
  if (!ok) throw Error{ECSPRT_ERR_ARG, path + " is shorter than its header says"};
  if (bs) {
    *bs = 1;
    if (FILE *g = std::fopen((path + ".info").c_str(), "r")) {
      int b = 0;
      if (std::fscanf(g, "-vecload_block_size %d", &b) == 1 && b > 0) *bs = b;
      std::fclose(g);
    }
  }
  return n;
}

// Grid shape of a context's DMDA vectors: [Nz][Ny][Nx], or [Ny][Nx] in two dimensions.
int grid_dims(const ecsprt_params &p, hsize_t *dims) {
  if (p.ndim == 2) {
    dims[0] = (hsize_t)p.Ny;
    dims[1] = (hsize_t)p.Nx;
    return 2;
  }
  dims[0] = (hsize_t)p.Nz;
  dims[1] = (hsize_t)p.Ny;
  dims[2] = (hsize_t)p.Nx;
  return 3;
}

// The fluid file: the four FluidNode components as "density", "x flux", "y flux",
// "z flux" (the field names of src/setup.c:388-391), then "rhs" and "potential".
void write_fluid(const std::string &path, int nd, const hsize_t *dims, long n, const double *moments,
                 const double *rhs, const double *phi) {
  std::vector<double> comp(4 * (size_t)n);
  for (long q = 0; q < n; ++q)
    for (int f = 0; f < 4; ++f) comp[(size_t)f * n + q] = moments[4 * q + f];
  h5_write(path, nd, dims, {"density", "x flux", "y flux", "z flux", "rhs", "potential"},
           {comp.data(), comp.data() + n, comp.data() + 2 * n, comp.data() + 3 * n, rhs, phi});
}

thread_local std::string g_io_error;

template <typename F>
int io_guard(F &&f) {
  try {
    f();
    return ECSPRT_OK;
  } catch (const Error &e) {
    g_io_error = e.msg;
    set_global_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    g_io_error = "host allocation failed";
    return ECSPRT_ERR_NOMEM;
  }
}

template <typename F>
int ctx_guard(ecsprt_hip_ctx *c, F &&f) {
  try {
    f();
    return ECSPRT_OK;
  } catch (const Error &e) {
    c->last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    c->last_error = "host allocation failed";
    return ECSPRT_ERR_NOMEM;
  }
}

}  // namespace
}  // namespace ecsprt

using namespace ecsprt;

extern "C" {

int ecsprt_io_hdf5_available(char *msg, int cap) {
  std::lock_guard<std::mutex> lock(g_h5_mu);
  Hdf5 &h = hdf5();
  if (msg && cap > 0) std::snprintf(msg, (size_t)cap, "%s", h.ok ? "" : h.why.c_str());
  return h.ok ? 1 : 0;
}

const char *ecsprt_io_last_error(void) { return g_io_error.c_str(); }

int ecsprt_io_write_fluid_hdf5(const char *path, int ndim, int Nx, int Ny, int Nz, const double *moments,
                               const double *rhs, const double *phi) {
  if (!path || !moments || !rhs || !phi || Nx < 1 || Ny < 1 || Nz < 1 || (ndim != 2 && ndim != 3))
    return ECSPRT_ERR_ARG;
  return io_guard([&] {
    hsize_t dims[3] = {(hsize_t)Nz, (hsize_t)Ny, (hsize_t)Nx};
    const int nd = ndim == 2 ? 2 : 3;
    if (nd == 2) {
      dims[0] = (hsize_t)Ny;
      dims[1] = (hsize_t)Nx;
    }
    write_fluid(path, nd, dims, (long)Nx * Ny * (ndim == 2 ? 1 : Nz), moments, rhs, phi);
  });
}

int ecsprt_io_write_hdf5_dataset(const char *path, const char *name, int ndims, const long *dims, const double *data) {
  if (!path || !name || !dims || !data || ndims < 1 || ndims > 4) return ECSPRT_ERR_ARG;
  return io_guard([&] {
    hsize_t d[4];
    for (int q = 0; q < ndims; ++q) {
      if (dims[q] < 1) throw Error{ECSPRT_ERR_ARG, "dataset dimensions must be positive"};
      d[q] = (hsize_t)dims[q];
    }
    h5_write(path, ndims, d, {name}, {data});
  });
}

int ecsprt_io_read_hdf5_dataset(const char *path, const char *name, int *ndims, long *dims, double *out, long cap) {
  if (!path || !name || !ndims || !dims) return ECSPRT_ERR_ARG;
  return io_guard([&] {
    hsize_t d[4] = {0, 0, 0, 0};
    *ndims = h5_read(path, name, d, out, cap);
    for (int q = 0; q < 4; ++q) dims[q] = q < *ndims ? (long)d[q] : 0;
  });
}

int ecsprt_io_write_vec_binary(const char *path, long n, int bs, const double *v) {
  if (!path || (!v && n > 0) || bs < 1) return ECSPRT_ERR_ARG;
  return io_guard([&] { vec_write(path, n, bs, v); });
}

int ecsprt_io_read_vec_binary(const char *path, double *v, long cap, long *n, int *bs) {
  if (!path || !n) return ECSPRT_ERR_ARG;
  return io_guard([&] { *n = vec_read(path, v, cap, bs); });
}

int ecsprt_hip_output_fluid_hdf5(ecsprt_hip_ctx *c, const char *insert) {
  if (!c) return ECSPRT_ERR_ARG;
  return ctx_guard(c, [&] {
    hsize_t dims[3];
    const int nd = grid_dims(c->p, dims);
    const long n = c->nnode;
    std::vector<double> m(4 * (size_t)n), rhs((size_t)n), phi((size_t)n);
    ECS_HIP(hipMemcpyAsync(m.data(), c->d_moments, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, c->stream));
    ECS_HIP(hipMemcpyAsync(rhs.data(), c->d_rhs, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    ECS_HIP(hipMemcpyAsync(phi.data(), c->d_phi, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
    write_fluid(std::string("fluid") + (insert ? insert : "") + ".hdf", nd, dims, n, m.data(), rhs.data(), phi.data());
  });
}

int ecsprt_hip_output_swarm_binary(ecsprt_hip_ctx *c, const char *insert) {
  if (!c) return ECSPRT_ERR_ARG;
  const std::string tag = insert ? insert : "";
  const long n = c->np;
  const int nd = c->p.ndim == 2 ? 2 : 3;
  std::vector<double> pos, vel;
  try {
    pos.resize((size_t)nd * (size_t)n);
    vel.resize((size_t)nd * (size_t)n);
  } catch (const std::bad_alloc &) {
    c->last_error = "host allocation failed";
    return ECSPRT_ERR_NOMEM;
  }
  const int rc = n > 0 ? ecsprt_hip_get_particles(c, pos.data(), vel.data()) : ECSPRT_OK;
  if (rc != ECSPRT_OK) return rc;
  return ctx_guard(c, [&] {
    vec_write("position" + tag + ".bin", nd * n, nd, pos.data());
    vec_write("velocity" + tag + ".bin", nd * n, nd, vel.data());
  });
}

int ecsprt_hip_load_fluid_quantities(ecsprt_hip_ctx *c, const double *flux_scale, const char *inpath) {
  if (!c || !flux_scale) return ECSPRT_ERR_ARG;
  return ctx_guard(c, [&] {
    if (!inpath || !inpath[0]) {
      c->rhs_sinusoidal = true;  // the reference's fallback to ComputeSinusoidalRHS3D
      return;
    }
    hsize_t want[3], got[4] = {0, 0, 0, 0};
    const int nd = grid_dims(c->p, want);
    const long n = c->nnode;
    if (h5_read(inpath, "density-kji", got, nullptr, 0) != nd || std::memcmp(want, got, sizeof(hsize_t) * nd) != 0)
      throw Error{ECSPRT_ERR_ARG, std::string("dataset \"density-kji\" in ") + inpath +
                                      " does not have the shape of the grid"};
    std::vector<double> dens((size_t)n), m(4 * (size_t)n);
    h5_read(inpath, "density-kji", got, dens.data(), n);
    for (long q = 0; q < n; ++q) {
      m[4 * q] = dens[q];
      for (int f = 0; f < 3; ++f) m[4 * q + 1 + f] = 0.0 + flux_scale[f] * dens[q];
    }
    ECS_HIP(hipMemcpyAsync(c->d_moments, m.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
    c->rhs_sinusoidal = false;
    c->operator_built = false;  // the next solve builds its operator from the loaded density
  });
}

}  // extern "C"
