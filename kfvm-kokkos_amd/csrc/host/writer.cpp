// Plot / checkpoint files compatible with the reference's PDI/HDF5 writer.
// See writer.hpp.
#include "writer.hpp"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <type_traits>

namespace kfvm {

namespace {

// The part of the HDF5 1.10+ C API the writer uses (hid_t is 64-bit there).
using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
constexpr hid_t P_DEFAULT = 0, S_ALL = 0;
constexpr unsigned F_ACC_TRUNC = 0x0002u, F_ACC_RDONLY = 0x0000u, F_ACC_RDWR = 0x0001u;
constexpr int S_SCALAR = 0, S_SELECT_SET = 0;

struct H5 {
  herr_t (*open)();
  herr_t (*get_libversion)(unsigned *, unsigned *, unsigned *);
  herr_t (*eset_auto)(hid_t, void *, void *);
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*screate)(int);
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
  herr_t (*sclose)(hid_t);
  int (*sget_ndims)(hid_t);
  int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
  herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *, const hsize_t *);
  hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen)(hid_t, const char *, hid_t);
  hid_t (*dget_space)(hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  herr_t (*dclose)(hid_t);
  int (*lexists)(hid_t, const char *, hid_t);
  hid_t *native_double, *native_int;
};

const H5 &h5() {
  static H5 api;
  static std::string err;
  static std::once_flag once;
  std::call_once(once, [] {
    std::vector<std::string> names;
    if (const char *e = std::getenv("KFVM_HDF5_LIB")) names.push_back(e);
    names.insert(names.end(), {"libhdf5.so", "libhdf5.so.310", "libhdf5.so.200", "libhdf5.so.103",
                               "/opt/conda/lib/libhdf5.so"});
    void *lib = nullptr;
    for (const std::string &n : names)
      if ((lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
    if (!lib) {
      err = "HDF5 library not found (set KFVM_HDF5_LIB to libhdf5.so)";
      return;
    }
    auto sym = [&](const char *s) {
      void *p = dlsym(lib, s);
      if (!p && err.empty()) err = std::string("HDF5 library lacks ") + s;
      return p;
    };
#define KFVM_H5SYM(field, name) api.field = reinterpret_cast<decltype(api.field)>(sym(name))
    KFVM_H5SYM(open, "H5open");
    KFVM_H5SYM(get_libversion, "H5get_libversion");
    KFVM_H5SYM(eset_auto, "H5Eset_auto2");
    KFVM_H5SYM(fcreate, "H5Fcreate");
    KFVM_H5SYM(fopen, "H5Fopen");
    KFVM_H5SYM(fclose, "H5Fclose");
    KFVM_H5SYM(screate, "H5Screate");
    KFVM_H5SYM(screate_simple, "H5Screate_simple");
    KFVM_H5SYM(sclose, "H5Sclose");
    KFVM_H5SYM(sget_ndims, "H5Sget_simple_extent_ndims");
    KFVM_H5SYM(sget_dims, "H5Sget_simple_extent_dims");
    KFVM_H5SYM(sselect_hyperslab, "H5Sselect_hyperslab");
    KFVM_H5SYM(dcreate, "H5Dcreate2");
    KFVM_H5SYM(dopen, "H5Dopen2");
    KFVM_H5SYM(dget_space, "H5Dget_space");
    KFVM_H5SYM(dwrite, "H5Dwrite");
    KFVM_H5SYM(dread, "H5Dread");
    KFVM_H5SYM(dclose, "H5Dclose");
    KFVM_H5SYM(lexists, "H5Lexists");
    KFVM_H5SYM(native_double, "H5T_NATIVE_DOUBLE_g");
    KFVM_H5SYM(native_int, "H5T_NATIVE_INT_g");
#undef KFVM_H5SYM
    if (!err.empty()) return;
    unsigned maj = 0, min = 0, rel = 0;
    api.get_libversion(&maj, &min, &rel);
    if (maj < 1 || (maj == 1 && min < 10)) {
      err = "HDF5 1.10 or newer required";
      return;
    }
    api.open();  // initialises the native type identifiers
    api.eset_auto(0, nullptr, nullptr);
  });
  if (!err.empty()) throw std::runtime_error(err);
  return api;
}

void chk(herr_t r, const char *what) {
  if (r < 0) throw std::runtime_error(std::string("HDF5: ") + what + " failed");
}
hid_t chkid(hid_t r, const char *what) {
  if (r < 0) throw std::runtime_error(std::string("HDF5: ") + what + " failed");
  return r;
}

void write_array(hid_t file, const char *name, const double *data, int rank, const hsize_t *dims) {
  const H5 &H = h5();
  const hid_t sp = chkid(H.screate_simple(rank, dims, nullptr), "create dataspace");
  const hid_t ds = H.dcreate(file, name, *H.native_double, sp, P_DEFAULT, P_DEFAULT, P_DEFAULT);
  if (ds < 0) {
    H.sclose(sp);
    throw std::runtime_error(std::string("HDF5: create dataset ") + name + " failed");
  }
  const herr_t r = H.dwrite(ds, *H.native_double, S_ALL, S_ALL, P_DEFAULT, data);
  H.dclose(ds);
  H.sclose(sp);
  chk(r, name);
}

template <class T>
void write_scalar(hid_t file, const char *name, T v) {
  const H5 &H = h5();
  const hid_t type = std::is_same<T, int>::value ? *H.native_int : *H.native_double;
  const hid_t sp = chkid(H.screate(S_SCALAR), "create scalar dataspace");
  const hid_t ds = H.dcreate(file, name, type, sp, P_DEFAULT, P_DEFAULT, P_DEFAULT);
  if (ds < 0) {
    H.sclose(sp);
    throw std::runtime_error(std::string("HDF5: create dataset ") + name + " failed");
  }
  const herr_t r = H.dwrite(ds, type, S_ALL, S_ALL, P_DEFAULT, &v);
  H.dclose(ds);
  H.sclose(sp);
  chk(r, name);
}

// Read dataset `name` of exactly n doubles (or ints, int_ = true).
void read_data(hid_t file, const char *name, void *out, size_t n, bool int_) {
  const H5 &H = h5();
  const hid_t ds = H.dopen(file, name, P_DEFAULT);
  if (ds < 0) throw std::runtime_error(std::string("checkpoint lacks dataset ") + name);
  const hid_t sp = H.dget_space(ds);
  const int rank = H.sget_ndims(sp);
  size_t count = 1;
  if (rank > 0) {
    std::vector<hsize_t> dims(static_cast<size_t>(rank));
    H.sget_dims(sp, dims.data(), nullptr);
    for (hsize_t d : dims) count *= size_t(d);
  }
  H.sclose(sp);
  if (count != n) {
    H.dclose(ds);
    throw std::runtime_error(std::string("checkpoint dataset ") + name + " has the wrong size");
  }
  const herr_t r = H.dread(ds, int_ ? *H.native_int : *H.native_double, S_ALL, S_ALL, P_DEFAULT, out);
  H.dclose(ds);
  chk(r, name);
}

std::string step_name(const std::string &base, int step, const char *ext) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "_%07d.%s", step, ext);
  return base + buf;
}

}  // namespace

std::vector<std::string> Writer::consNames(Eq eq) {
  if (eq == Eq::LinAdv) return {"dens"};
  if (eq == Eq::MhdGlm) return {"dens", "momx", "momy", "momz", "etot", "magx", "magy", "magz", "psi"};
  return {"dens", "momx", "momy", "momz", "etot"};
}

Writer::Writer(const Config &cfg) : cfg_(cfg), nv_(num_vars(cfg.eq)) {
  nx_ = cfg.nX;
  ny_ = cfg.nY;
  nz_ = cfg.dim == 3 ? cfg.nZ : 1;
  std::ostringstream o;
  o << cfg.dataDir << "/R" << cfg.rad << "_NX" << nx_ << "_NY" << ny_;
  if (cfg.dim == 3) o << "_NZ" << nz_;
  o << "/";
  prefix_ = o.str();
  // nodal coordinates, z even in 2D (spacing dmin there), WriterPDI.cpp:38-50
  const double dz = cfg.dim == 2 ? cfg.dmin() : cfg.dz();
  for (int64_t n = 0; n <= nx_; n++) xc_.push_back(double(n) * cfg.dx() + cfg.xLo);
  for (int64_t n = 0; n <= ny_; n++) yc_.push_back(double(n) * cfg.dy() + cfg.yLo);
  for (int64_t n = 0; n <= nz_; n++) zc_.push_back(double(n) * dz + cfg.zLo);
}

void Writer::writeFile(const std::string &path, const OutputFields &f, bool plot, int step, double time,
                       double dt, double errEst) {
  const H5 &H = h5();
  const hid_t file = H.fcreate(path.c_str(), F_ACC_TRUNC, P_DEFAULT, P_DEFAULT);
  if (file < 0) throw std::runtime_error("cannot create " + path);
  try {
    const size_t nc = size_t(nx_) * ny_ * nz_;
    const hsize_t dims[3] = {hsize_t(nz_), hsize_t(ny_), hsize_t(nx_)};
    const std::vector<std::string> names = consNames(cfg_.eq);
    for (int v = 0; v < nv_; v++) write_array(file, names[size_t(v)].c_str(), f.U + size_t(v) * nc, 3, dims);
    write_array(file, "weno", f.flags, 3, dims);
    write_array(file, "fomx", f.flags + nc, 3, dims);
    if (plot && cfg_.eq != Eq::LinAdv) {
      // auxiliary variables (AuxVars, Hydro_K.H:66-100 / MHD_GLM_K.H:72-106)
      const bool mhd = cfg_.eq == Eq::MhdGlm;
      const int na = mhd ? 6 : 5;
      std::vector<double> V(size_t(na) * nc);
      const double gm1 = cfg_.gamma - 1.0;
      for (size_t c = 0; c < nc; c++) {
        double u[9];
        for (int v = 0; v < nv_; v++) u[v] = f.U[size_t(v) * nc + c];
        for (int d = 0; d < 3; d++) V[size_t(d) * nc + c] = u[1 + d] / u[0];
        const double ekin = 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0];
        if (mhd) {
          const double prsb = 0.5 * (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]);
          const double eint = u[4] - prsb - ekin - 0.5 * u[8] * u[8];
          V[3 * nc + c] = eint;
          V[4 * nc + c] = gm1 * eint;
          V[5 * nc + c] = prsb;
        } else {
          const double eint = u[4] - ekin;
          V[3 * nc + c] = eint;
          V[4 * nc + c] = gm1 * eint;
        }
      }
      const char *an[6] = {"velx", "vely", "velz", "eint", mhd ? "prsg" : "pres", "prsb"};
      for (int a = 0; a < na; a++) write_array(file, an[a], V.data() + size_t(a) * nc, 3, dims);
    }
    const hsize_t nxd = hsize_t(xc_.size()), nyd = hsize_t(yc_.size()), nzd = hsize_t(zc_.size());
    write_array(file, "xcoord", xc_.data(), 1, &nxd);
    write_array(file, "ycoord", yc_.data(), 1, &nyd);
    write_array(file, "zcoord", zc_.data(), 1, &nzd);
    write_scalar(file, "time", time);
    write_scalar(file, "time_step", step);
    write_scalar(file, "gamma", cfg_.gamma);
    if (!plot) {
      write_scalar(file, "time_step_size", dt);
      // not in the reference's checkpoints: restores the controller exactly
      write_scalar(file, "errEst", errEst);
    }
  } catch (...) {
    H.fclose(file);
    throw;
  }
  chk(H.fclose(file), "close file");
}

void Writer::writeXdmf(const std::string &path, const std::string &h5name, double time, bool plot) const {
  std::ofstream o(path, std::ios::trunc);
  if (!o) throw std::runtime_error("cannot create " + path);
  o.precision(17);
  auto scalar = [&](const char *n) {
    o << "      <Attribute Name=\"" << n << "\" AttributeType=\"Scalar\" Center=\"Cell\">\n"
      << "        <DataItem Format=\"HDF\" NumberType=\"Float\" Precision=\"8\" Dimensions=\"" << nz_ << " "
      << ny_ << " " << nx_ << "\">\n          " << h5name << ":/" << n << "\n        </DataItem>\n"
      << "      </Attribute>\n\n";
  };
  auto vector = [&](const char *n, const char *a, const char *b, const char *c) {
    o << "      <Attribute Name=\"" << n << "\" AttributeType=\"Vector\" Center=\"Cell\">\n"
      << "        <DataItem ItemType=\"Function\" Function=\"JOIN($0, $1, $2)\" Dimensions=\"" << nz_ << " "
      << ny_ << " " << nx_ << " 3\">\n";
    for (const char *x : {a, b, c})
      o << "          <DataItem Format=\"HDF\" NumberType=\"Float\" Precision=\"8\" Dimensions=\"" << nz_ << " "
        << ny_ << " " << nx_ << "\">\n            " << h5name << ":/" << x << "\n          </DataItem>\n";
    o << "        </DataItem>\n      </Attribute>\n\n";
  };
  o << "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n<Xdmf Version=\"2.0\">\n  <Domain>\n"
    << "    <Grid Name=\"Structured Grid\" GridType=\"Uniform\">\n      <Time Value=\"" << time << "\" />\n"
    << "      <Topology TopologyType=\"3DRectMesh\" NumberOfElements=\"" << (nz_ + 1) << " " << (ny_ + 1)
    << " " << (nx_ + 1) << "\"/>\n      <Geometry GeometryType=\"VxVyVz\">\n";
  const char *axn[3] = {"Vx", "Vy", "Vz"}, *cn[3] = {"xcoord", "ycoord", "zcoord"};
  const int64_t ext[3] = {nx_ + 1, ny_ + 1, nz_ + 1};
  for (int a = 0; a < 3; a++)
    o << "        <DataItem Name=\"" << axn[a] << "\" Dimensions=\"" << ext[a]
      << "\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">\n          " << h5name << ":/" << cn[a]
      << "\n        </DataItem>\n";
  o << "      </Geometry>\n\n";
  if (cfg_.eq == Eq::LinAdv) {
    scalar("dens");
  } else if (cfg_.eq == Eq::MhdGlm) {
    scalar("dens");
    vector("mom", "momx", "momy", "momz");
    vector("B", "magx", "magy", "magz");
    scalar("etot");
    scalar("psi");
    if (plot) {
      vector("vel", "velx", "vely", "velz");
      scalar("eint");
      scalar("prsg");
      scalar("prsb");
    }
  } else {
    scalar("dens");
    vector("mom", "momx", "momy", "momz");
    scalar("etot");
    if (plot) {
      vector("vel", "velx", "vely", "velz");
      scalar("eint");
      scalar("pres");
    }
  }
  scalar("weno");
  scalar("fomx");
  o << "    </Grid>\n  </Domain>\n</Xdmf>\n";
}

std::string Writer::writePlot(const OutputFields &f, int num, double time) {
  std::filesystem::create_directories(prefix_ + "plot/");
  const std::string h5name = step_name(cfg_.baseName, num, "h5");
  writeXdmf(prefix_ + "plot/" + step_name(cfg_.baseName, num, "xmf"), h5name, time, true);
  writeFile(prefix_ + "plot/" + h5name, f, true, num, time, 0.0, 1.0);
  return prefix_ + "plot/" + h5name;
}

// Flow statistics file (WriterPDI::writeFlowStats, WriterPDI.cpp:69-86 and
// the stats_data dataset of user/pdi_*_xdmf.yml): <prefix><basename>_stats.h5
// holds stats_data [maxRows][1 + nstats] (time, then the statistics), row
// `row` written in place; the file is created (zero-filled) with the first
// row.
void Writer::writeStats(int row, int maxRows, double time, const double *stats, int nstats) {
  const H5 &H = h5();
  if (row < 0 || row >= maxRows) throw std::runtime_error("flow statistics: row beyond time:maxsteps");
  std::filesystem::create_directories(prefix_);
  const std::string path = prefix_ + cfg_.baseName + "_stats.h5";
  const hsize_t ncol = hsize_t(nstats + 1);
  hid_t file, ds;
  if (row == 0) {
    file = chkid(H.fcreate(path.c_str(), F_ACC_TRUNC, P_DEFAULT, P_DEFAULT), "create stats file");
    const hsize_t dims[2] = {hsize_t(maxRows), ncol};
    const hid_t sp = chkid(H.screate_simple(2, dims, nullptr), "create dataspace");
    ds = H.dcreate(file, "stats_data", *H.native_double, sp, P_DEFAULT, P_DEFAULT, P_DEFAULT);
    H.sclose(sp);
  } else {
    file = chkid(H.fopen(path.c_str(), F_ACC_RDWR, P_DEFAULT), "open stats file");
    ds = H.dopen(file, "stats_data", P_DEFAULT);
  }
  if (ds < 0) {
    H.fclose(file);
    throw std::runtime_error("HDF5: stats_data dataset failed");
  }
  std::vector<double> rowv(static_cast<size_t>(ncol));
  rowv[0] = time;
  for (int n = 0; n < nstats; n++) rowv[size_t(n) + 1] = stats[n];
  const hid_t fsp = H.dget_space(ds);
  const hsize_t start[2] = {hsize_t(row), 0}, count[2] = {1, ncol};
  herr_t r = H.sselect_hyperslab(fsp, S_SELECT_SET, start, nullptr, count, nullptr);
  const hid_t msp = H.screate_simple(2, count, nullptr);
  if (r >= 0) r = H.dwrite(ds, *H.native_double, msp, fsp, P_DEFAULT, rowv.data());
  H.sclose(msp);
  H.sclose(fsp);
  H.dclose(ds);
  H.fclose(file);
  chk(r, "write stats row");
}

std::string Writer::writeCkpt(const OutputFields &f, int step, double time, double dt, double errEst) {
  std::filesystem::create_directories(prefix_ + "ckpt/");
  const std::string h5name = step_name(cfg_.baseName, step, "h5");
  writeXdmf(prefix_ + "ckpt/" + step_name(cfg_.baseName, step, "xmf"), h5name, time, false);
  writeFile(prefix_ + "ckpt/" + h5name, f, false, step, time, dt, errEst);
  return prefix_ + "ckpt/" + h5name;
}

CkptData Writer::readCkpt(const std::string &path) const {
  const H5 &H = h5();
  const hid_t file = H.fopen(path.c_str(), F_ACC_RDONLY, P_DEFAULT);
  if (file < 0) throw std::runtime_error("cannot open checkpoint " + path);
  CkptData d;
  try {
    const size_t nc = size_t(nx_) * ny_ * nz_;
    d.U.resize(size_t(nv_) * nc);
    const std::vector<std::string> names = consNames(cfg_.eq);
    for (int v = 0; v < nv_; v++) read_data(file, names[size_t(v)].c_str(), d.U.data() + size_t(v) * nc, nc, false);
    read_data(file, "time", &d.time, 1, false);
    read_data(file, "time_step", &d.step, 1, true);
    read_data(file, "time_step_size", &d.dt, 1, false);
    if (H.lexists(file, "fomx", P_DEFAULT) > 0) {
      d.theta.resize(nc);
      read_data(file, "fomx", d.theta.data(), nc, false);
    }
    if (H.lexists(file, "errEst", P_DEFAULT) > 0) {
      read_data(file, "errEst", &d.errEst, 1, false);
      d.haveErrEst = true;
    }
  } catch (...) {
    H.fclose(file);
    throw;
  }
  H.fclose(file);
  return d;
}

}  // namespace kfvm
