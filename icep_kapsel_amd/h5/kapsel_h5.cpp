// kapsel_h5.cpp -- HDF5 field and particle dumps in the reference's
// hdf5_writer layout (file format in include/kapsel_h5.h).  Host code only,
// against the HDF5 1.10 C API.
#include "../../include/kapsel_h5.h"

#include <hdf5.h>
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kapsel_gpu.h"

namespace {

thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

struct H5Error {};

hid_t check(hid_t id, const char *what) {
  if (id < 0) {
    set_error("HDF5: %s failed", what);
    throw H5Error{};
  }
  return id;
}

// attribute of a scalar type on an object
void put_attr(hid_t loc, const char *name, hid_t type, const void *data, hsize_t n) {
  hid_t space = check(H5Screate_simple(1, &n, nullptr), "H5Screate_simple");
  hid_t attr = H5Acreate2(loc, name, type, space, H5P_DEFAULT, H5P_DEFAULT);
  if (attr < 0) {
    H5Sclose(space);
    check(attr, "H5Acreate2");
  }
  herr_t st = H5Awrite(attr, type, data);
  H5Aclose(attr);
  H5Sclose(space);
  check(st, "H5Awrite");
}

void put_dataset(hid_t loc, const char *name, hid_t mem_type, hid_t file_type, const void *data, int rank,
                 const hsize_t *dims) {
  hid_t space = check(H5Screate_simple(rank, dims, nullptr), "H5Screate_simple");
  hid_t ds = H5Dcreate2(loc, name, file_type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (ds < 0) {
    H5Sclose(space);
    set_error("HDF5: cannot create dataset %s", name);
    throw H5Error{};
  }
  herr_t st = H5Dwrite(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  H5Dclose(ds);
  H5Sclose(space);
  check(st, "H5Dwrite");
}

}  // namespace

struct kapsel_h5 {
  int nx, ny, nz, nz_;
  float dx, dt;
  int nump_total;
  std::vector<int> plist;
  kapsel_h5_crop crop;
  bool fields;
  std::string dir, name, path;
  hid_t fid = -1, gid_sys = -1, gid_trj = -1, gid_time = -1, gid_field = -1, gid_part = -1;
  int ts = 0;
  std::vector<double> buf;

  // the cropped grid of a padded field, in i-j-k order (the reference's
  // element selection of the padded memory space, output_writer.cxx:97-123)
  const double *crop_of(const double *a) {
    buf.resize((size_t)crop.count[0] * crop.count[1] * crop.count[2]);
    size_t q = 0;
    for (int ii = 0; ii < crop.count[0]; ii++) {
      const int i = crop.start[0] + ii * crop.stride[0];
      for (int jj = 0; jj < crop.count[1]; jj++) {
        const int j = crop.start[1] + jj * crop.stride[1];
        for (int kk = 0; kk < crop.count[2]; kk++) {
          const int k = crop.start[2] + kk * crop.stride[2];
          buf[q++] = a[((size_t)i * ny + j) * nz_ + k];
        }
      }
    }
    return buf.data();
  }
  void field_dataset(hid_t loc, const char *nm, const double *padded) {
    const hsize_t dims[3] = {(hsize_t)crop.count[0], (hsize_t)crop.count[1], (hsize_t)crop.count[2]};
    put_dataset(loc, nm, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, crop_of(padded), 3, dims);
  }
  void vectors(const char *nm, const std::vector<double> &v) {
    const hsize_t dims[2] = {(hsize_t)plist.size(), 3};
    put_dataset(gid_part, nm, H5T_NATIVE_DOUBLE, H5T_NATIVE_FLOAT, v.data(), 2, dims);
  }
  void write_conf() const {
    const std::string fn = path + ".conf";
    FILE *f = fopen(fn.c_str(), "w");
    if (!f) {
      set_error("cannot write %s", fn.c_str());
      throw H5Error{};
    }
    const int np = (int)plist.size();
    fprintf(f, "#Configure file needed to generate xdmf trajectory\n#v1.1\n");
    fprintf(f, "[Filename]\n %s\n", name.c_str());
    fprintf(f, "[Grid Dimensions]\n %d %d %d\n", crop.count[0], crop.count[1], crop.count[2]);
    fprintf(f, "[Time Series]\n 0.0 %.6f %d\n", dt, ts);
    fprintf(f, "[Particle Number]\n %d\n[Obstacle Number]\n 0\n\n", np);
    if (np > 0) {
      fprintf(f, "[Particle Grid]\n pos\n#[Particle Grid]\n#pos_raw\n");
      fprintf(f, "[Particle Scalar Const]\n id\n[Particle Scalar Const]\n spec\n");
      for (const char *v : {"vel", "omega", "force_h", "force_r", "torque_h", "torque_r"})
        fprintf(f, "[Particle Vector]\n %s\n", v);
      fprintf(f, "[Particle Tensor]\n QR\n\n");
    }
    if (fields) {
      fprintf(f, "[Field GridX Const]\n x\n[Field GridY Const]\n y\n[Field GridZ Const]\n z\n");
      fprintf(f, "[Field Vector]\n u_x u_y u_z u\n[Field Scalar]\n phi\n\n");
    }
    fclose(f);
  }
};

extern "C" {

const char *kapsel_h5_last_error(void) { return g_err; }

int kapsel_h5_create(const char *out_dir, const char *out_name, int nx, int ny, int nz, double dx, double dt,
                     int nump_total, int nump_print, const int *print_list, const int *spec,
                     const kapsel_h5_crop *crop, int with_fields, kapsel_h5 **out) {
  if (!out_dir || !out_name || !out || nx < 1 || ny < 1 || nz < 1 || !(dx > 0) || nump_total < 0 ||
      nump_print < 0 || nump_print > nump_total) {
    set_error("kapsel_h5_create: invalid argument");
    return KAPSEL_ERR_ARG;
  }
  kapsel_h5 *h = new kapsel_h5();
  try {
    h->nx = nx;
    h->ny = ny;
    h->nz = nz;
    h->nz_ = 2 * (nz / 2 + 1);
    h->dx = (float)dx;
    h->dt = (float)dt;
    h->nump_total = nump_total;
    for (int j = 0; j < nump_print; j++) {
      const int i = print_list ? print_list[j] : j;
      if (i < 0 || i >= nump_total) {
        set_error("kapsel_h5_create: particle id %d out of range", i);
        throw H5Error{};
      }
      h->plist.push_back(i);
    }
    if (crop) {
      h->crop = *crop;
      const int n[3] = {nx, ny, nz};
      for (int d = 0; d < 3; d++)
        if (crop->count[d] < 1 || crop->stride[d] < 1 || crop->start[d] < 0 ||
            crop->start[d] + (crop->count[d] - 1) * crop->stride[d] >= n[d]) {
          set_error("kapsel_h5_create: crop outside the grid along axis %d", d);
          throw H5Error{};
        }
    } else {
      h->crop = kapsel_h5_crop{{0, 0, 0}, {nx, ny, nz}, {1, 1, 1}};
    }
    h->fields = with_fields != 0;
    h->dir = out_dir;
    h->name = out_name;
    h->path = h->dir + "/" + h->name;
    if (mkdir(out_dir, 0755) != 0 && errno != EEXIST) {
      set_error("kapsel_h5_create: cannot make %s", out_dir);
      throw H5Error{};
    }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    const std::string fn = h->path + ".h5";
    h->fid = check(H5Fcreate(fn.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate");
    const hid_t root = h->fid;
    put_attr(root, "dx", H5T_NATIVE_FLOAT, &h->dx, 1);
    const long long dims[3] = {h->crop.count[0], h->crop.count[1], h->crop.count[2]};
    put_attr(root, "nxnynz", H5T_NATIVE_LLONG, dims, 3);
    const float origin[3] = {h->crop.start[0] * h->dx, h->crop.start[1] * h->dx, h->crop.start[2] * h->dx};
    put_attr(root, "origin", H5T_NATIVE_FLOAT, origin, 3);
    const int np = (int)h->plist.size(), nobs = 0;
    put_attr(root, "nump", H5T_NATIVE_INT, &np, 1);
    put_attr(root, "nump_obs", H5T_NATIVE_INT, &nobs, 1);
    put_attr(root, "nump_total", H5T_NATIVE_INT, &h->nump_total, 1);
    h->gid_sys = check(H5Gcreate2(h->fid, "system_data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2");
    h->gid_trj = check(H5Gcreate2(h->fid, "trajectory_data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2");
    if (h->fields) {
      // write_field_info (output_writer.cxx:666-695): the grid coordinates
      hid_t g = check(H5Gcreate2(h->gid_sys, "./field", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2");
      std::vector<double> c((size_t)nx * ny * h->nz_, 0.0);
      for (int d = 0; d < 3; d++) {
        for (int i = 0; i < nx; i++)
          for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
              c[((size_t)i * ny + j) * h->nz_ + k] = (double)(d == 0 ? i : d == 1 ? j : k) * (double)h->dx;
        h->field_dataset(g, d == 0 ? "x" : d == 1 ? "y" : "z", c.data());
      }
      H5Gclose(g);
    }
    if (np > 0) {
      // write_particle_info (output_writer.cxx:697-726): ids and species
      hid_t g = check(H5Gcreate2(h->gid_sys, "./particle", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2");
      std::vector<int> sp(np);
      for (int j = 0; j < np; j++) sp[j] = spec ? spec[h->plist[j]] : 0;
      const hsize_t n1 = (hsize_t)np;
      put_dataset(g, "id", H5T_NATIVE_INT, H5T_NATIVE_INT, h->plist.data(), 1, &n1);
      put_dataset(g, "spec", H5T_NATIVE_INT, H5T_NATIVE_INT, sp.data(), 1, &n1);
      H5Gclose(g);
    }
    check(H5Fflush(h->fid, H5F_SCOPE_LOCAL), "H5Fflush");
    h->write_conf();
  } catch (const H5Error &) {
    if (h->gid_sys >= 0) H5Gclose(h->gid_sys);
    if (h->gid_trj >= 0) H5Gclose(h->gid_trj);
    if (h->fid >= 0) H5Fclose(h->fid);
    delete h;
    return KAPSEL_ERR_STATE;
  }
  *out = h;
  return KAPSEL_OK;
}

int kapsel_h5_frame_begin(kapsel_h5 *h) {
  if (!h || h->gid_time >= 0) {
    set_error("kapsel_h5_frame_begin: no writer, or a frame is open");
    return KAPSEL_ERR_STATE;
  }
  try {
    char nm[64];
    snprintf(nm, sizeof nm, "./frame_%d", h->ts);
    h->gid_time = check(H5Gcreate2(h->gid_trj, nm, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2");
    if (h->fields)
      h->gid_field = check(H5Gcreate2(h->gid_time, "./field", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2");
    if (!h->plist.empty())
      h->gid_part = check(H5Gcreate2(h->gid_time, "./particle", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2");
    const float t = (float)h->ts * h->dt;
    put_attr(h->gid_time, "time", H5T_NATIVE_FLOAT, &t, 1);
  } catch (const H5Error &) {
    return KAPSEL_ERR_STATE;
  }
  return KAPSEL_OK;
}

int kapsel_h5_write_field(kapsel_h5 *h, const char *name, const double *field) {
  if (!h || !name || !field || h->gid_field < 0) {
    set_error("kapsel_h5_write_field: no open frame with fields, or null argument");
    return KAPSEL_ERR_STATE;
  }
  try {
    h->field_dataset(h->gid_field, name, field);
  } catch (const H5Error &) {
    return KAPSEL_ERR_STATE;
  }
  return KAPSEL_OK;
}

int kapsel_h5_write_particles(kapsel_h5 *h, const double *state, const double *orientations) {
  if (!h || !state || h->gid_part < 0) {
    set_error("kapsel_h5_write_particles: no open frame with particles, or null argument");
    return KAPSEL_ERR_STATE;
  }
  const int n = h->nump_total, np = (int)h->plist.size();
  auto pv = [&](int f, int d, int i) { return state[((size_t)f * 3 + d) * n + i]; };
  try {
    std::vector<double> v((size_t)np * 3);
    const struct {
      const char *name;
      int field;
    } rows[] = {{"pos", KAPSEL_P_X},
                {"pos_raw", KAPSEL_P_X_NOPBC},
                {"vel", KAPSEL_P_V},
                {"force_h", KAPSEL_P_F_HYDRO_PREVIOUS},
                {"force_r", KAPSEL_P_FR_PREVIOUS},
                {"omega", KAPSEL_P_OMEGA},
                {"torque_h", KAPSEL_P_TORQUE_HYDRO_PREVIOUS},
                {"torque_r", KAPSEL_P_TORQUE_R_PREVIOUS}};
    for (const auto &r : rows) {
      for (int j = 0; j < np; j++)
        for (int d = 0; d < 3; d++) v[3 * j + d] = pv(r.field, d, h->plist[j]);
      h->vectors(r.name, v);
    }
    // QR: the rotation matrix R of the normalised orientation quaternion
    // q = (s, x, y, z), R = (1 - 2(yy + zz), 2(xy - sz), 2(xz + sy); ...), stored
    // with row j holding column j of R (the body axes in space)
    std::vector<double> qr((size_t)np * 9);
    for (int j = 0; j < np; j++) {
      const int i = h->plist[j];
      double q[4] = {1.0, 0.0, 0.0, 0.0};
      if (orientations)
        for (int c = 0; c < 4; c++) q[c] = orientations[(size_t)c * n + i];
      const double norm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
      for (double &c : q) c /= norm;
      const double s = q[0], x = q[1], y = q[2], z = q[3];
      const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - s * z), 2 * (x * z + s * y)},
                              {2 * (x * y + s * z), 1 - 2 * (x * x + z * z), 2 * (y * z - s * x)},
                              {2 * (x * z - s * y), 2 * (y * z + s * x), 1 - 2 * (x * x + y * y)}};
      for (int col = 0; col < 3; col++)
        for (int row = 0; row < 3; row++) qr[(size_t)9 * j + 3 * col + row] = R[row][col];
    }
    const hsize_t dims[3] = {(hsize_t)np, 3, 3};
    put_dataset(h->gid_part, "QR", H5T_NATIVE_DOUBLE, H5T_NATIVE_FLOAT, qr.data(), 3, dims);
  } catch (const H5Error &) {
    return KAPSEL_ERR_STATE;
  }
  return KAPSEL_OK;
}

int kapsel_h5_frame_end(kapsel_h5 *h) {
  if (!h || h->gid_time < 0) {
    set_error("kapsel_h5_frame_end: no open frame");
    return KAPSEL_ERR_STATE;
  }
  if (h->gid_part >= 0) H5Gclose(h->gid_part);
  if (h->gid_field >= 0) H5Gclose(h->gid_field);
  h->gid_part = h->gid_field = -1;
  herr_t st = H5Fflush(h->gid_time, H5F_SCOPE_LOCAL);
  H5Gclose(h->gid_time);
  h->gid_time = -1;
  h->ts += 1;
  if (st < 0) {
    set_error("HDF5: H5Fflush failed");
    return KAPSEL_ERR_STATE;
  }
  return KAPSEL_OK;
}

int kapsel_h5_close(kapsel_h5 *h) {
  if (!h) return KAPSEL_OK;
  int rc = KAPSEL_OK;
  if (h->gid_time >= 0) rc = kapsel_h5_frame_end(h);
  try {
    h->write_conf();
  } catch (const H5Error &) {
    rc = KAPSEL_ERR_STATE;
  }
  H5Gclose(h->gid_sys);
  H5Gclose(h->gid_trj);
  if (H5Fclose(h->fid) < 0) {
    set_error("HDF5: H5Fclose failed");
    rc = KAPSEL_ERR_STATE;
  }
  delete h;
  return rc;
}

int kapsel_h5_read(const char *path, const char *object, double *out, long long capacity, long long *count) {
  if (!path || !object || !count) {
    set_error("kapsel_h5_read: null argument");
    return KAPSEL_ERR_ARG;
  }
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  hid_t fid = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (fid < 0) {
    set_error("kapsel_h5_read: cannot open %s", path);
    return KAPSEL_ERR_ARG;
  }
  std::string obj(object);
  const size_t at = obj.find('@');
  int rc = KAPSEL_OK;
  hid_t id = -1, space = -1;
  bool attr = at != std::string::npos;
  if (attr) {
    const std::string owner = obj.substr(0, at), name = obj.substr(at + 1);
    id = H5Aopen_by_name(fid, owner.empty() ? "/" : owner.c_str(), name.c_str(), H5P_DEFAULT, H5P_DEFAULT);
    if (id >= 0) space = H5Aget_space(id);
  } else {
    id = H5Dopen2(fid, obj.c_str(), H5P_DEFAULT);
    if (id >= 0) space = H5Dget_space(id);
  }
  if (id < 0 || space < 0) {
    set_error("kapsel_h5_read: no %s %s in %s", attr ? "attribute" : "dataset", object, path);
    rc = KAPSEL_ERR_ARG;
  } else {
    const long long n = (long long)H5Sget_simple_extent_npoints(space);
    *count = n;
    if (out) {
      if (n > capacity) {
        set_error("kapsel_h5_read: %lld values, room for %lld", n, capacity);
        rc = KAPSEL_ERR_ARG;
      } else {
        herr_t st = attr ? H5Aread(id, H5T_NATIVE_DOUBLE, out)
                         : H5Dread(id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
        if (st < 0) {
          set_error("kapsel_h5_read: read of %s failed", object);
          rc = KAPSEL_ERR_STATE;
        }
      }
    }
  }
  if (space >= 0) H5Sclose(space);
  if (id >= 0) attr ? H5Aclose(id) : H5Dclose(id);
  H5Fclose(fid);
  return rc;
}

}  // extern "C"
