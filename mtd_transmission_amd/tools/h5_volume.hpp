// Complex index volume as HDF5, for the polarisation mode of
// tomo_reconstruction (save_complex_volume_hdf5,
// Reconstruction/src/IO_functions.cpp:52-84, called at
// Reconstruction/main.cpp:360-363): file <name>, group "/indice_complexe",
// datasets "Re" and "Im", float32, dimensions {Z, Y, X}.
//
// The HDF5 C library is loaded at run time (the image's libhdf5, HDF5 1.10
// ABI: hid_t is 64-bit), so neither libmtd.so nor the executables link it;
// write_complex_volume_h5 returns false when it cannot be loaded.
#pragma once
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <string>

namespace tomo {

namespace h5 {
using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kDefault = 0;     // H5P_DEFAULT, H5S_ALL
constexpr unsigned kAccTrunc = 2;  // H5F_ACC_TRUNC

struct Api {
  void* lib = nullptr;
  herr_t (*open)();
  hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
  hid_t (*gcreate2)(hid_t, const char*, hid_t, hid_t, hid_t);
  hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
  hid_t (*dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
  herr_t (*dclose)(hid_t);
  herr_t (*sclose)(hid_t);
  herr_t (*gclose)(hid_t);
  herr_t (*fclose)(hid_t);
  const hid_t* native_float;  // H5T_NATIVE_FLOAT_g, valid after open()

  bool load() {
    const char* env = std::getenv("MTD_HDF5_LIB");
    const char* names[] = {env, "libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
                           "/opt/conda/lib/libhdf5.so"};
    for (const char* n : names) {
      if (!n) continue;
      lib = dlopen(n, RTLD_NOW | RTLD_LOCAL);
      if (lib) break;
    }
    if (!lib) return false;
    auto sym = [this](const char* s) { return dlsym(lib, s); };
    open = (herr_t(*)())sym("H5open");
    fcreate = (hid_t(*)(const char*, unsigned, hid_t, hid_t))sym("H5Fcreate");
    gcreate2 = (hid_t(*)(hid_t, const char*, hid_t, hid_t, hid_t))sym("H5Gcreate2");
    screate_simple = (hid_t(*)(int, const hsize_t*, const hsize_t*))sym("H5Screate_simple");
    dcreate2 = (hid_t(*)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t))sym("H5Dcreate2");
    dwrite = (herr_t(*)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*))sym("H5Dwrite");
    dclose = (herr_t(*)(hid_t))sym("H5Dclose");
    sclose = (herr_t(*)(hid_t))sym("H5Sclose");
    gclose = (herr_t(*)(hid_t))sym("H5Gclose");
    fclose = (herr_t(*)(hid_t))sym("H5Fclose");
    native_float = (const hid_t*)sym("H5T_NATIVE_FLOAT_g");
    return open && fcreate && gcreate2 && screate_simple && dcreate2 && dwrite && dclose &&
           sclose && gclose && fclose && native_float && open() >= 0;
  }
};
}  // namespace h5

inline bool write_complex_volume_h5(const std::string& path, const float* re, const float* im,
                                    size_t nx, size_t ny, size_t nz) {
  static h5::Api api;
  static const bool ok = api.load();
  if (!ok) return false;
  const h5::hid_t f = api.fcreate(path.c_str(), h5::kAccTrunc, h5::kDefault, h5::kDefault);
  if (f < 0) return false;
  bool good = false;
  const h5::hid_t grp = api.gcreate2(f, "/indice_complexe", h5::kDefault, h5::kDefault,
                                     h5::kDefault);
  if (grp >= 0) {
    const h5::hsize_t dims[3] = {nz, ny, nx};
    const h5::hid_t sp = api.screate_simple(3, dims, nullptr);
    if (sp >= 0) {
      good = true;
      const char* names[2] = {"Re", "Im"};
      const float* data[2] = {re, im};
      for (int k = 0; k < 2 && good; ++k) {
        const h5::hid_t d = api.dcreate2(grp, names[k], *api.native_float, sp, h5::kDefault,
                                         h5::kDefault, h5::kDefault);
        good = d >= 0 && api.dwrite(d, *api.native_float, h5::kDefault, h5::kDefault,
                                    h5::kDefault, data[k]) >= 0;
        if (d >= 0) api.dclose(d);
      }
      api.sclose(sp);
    }
    api.gclose(grp);
  }
  return api.fclose(f) >= 0 && good;
}

}  // namespace tomo
