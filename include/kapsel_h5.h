/*
 * kapsel_h5.h -- C ABI of the HDF5 field and particle dumps
 * (icep_kapsel_amd/libkapsel_h5.so, host only, over libhdf5 1.10).
 *
 * The file is the reference's hdf5_writer output (output_writer.cxx:41-575):
 *   <out_dir>/<out_name>.h5 with root attributes dx (float), nxnynz
 *   (long long[3]), origin (float[3]), nump, nump_obs, nump_total (int);
 *   /system_data/field/{x,y,z} (the grid coordinates, double, cropped);
 *   /system_data/particle/{id,spec} (int);
 *   /trajectory_data/frame_<n>/ with attribute time = n*dt (float),
 *     field/<name> (double, the cropped NX x NY x NZ grid of a padded
 *     field) and particle/{pos,pos_raw,vel,force_h,force_r,omega,torque_h,
 *     torque_r} (float, N x 3) and particle/QR (float, N x 3 x 3, the
 *     body axes as rows: QR[j][0..2] = column j of the rotation matrix);
 *   and <out_dir>/<out_name>.conf, the xdmf configuration text
 *   (write_configure_file, output_writer.cxx:578-664).
 * Field names are the reference's: u_x u_y u_z phi pressure surface_charge
 * solute_charge e_potential plus_charge minus_charge epsilon f_maxwell_x ...
 * Host buffers use the padded layout of kapsel_gpu.h (NX*NY*NZ_ doubles) and
 * the particle state / orientation layouts of kapsel_get_particles /
 * kapsel_get_orientations.  Every function returns 0 or a negative code;
 * kapsel_h5_last_error() gives the message.
 */
#ifndef KAPSEL_H5_H
#define KAPSEL_H5_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kapsel_h5 kapsel_h5;

/* Field_crop (variable.h): start, count, stride per axis */
typedef struct kapsel_h5_crop {
  int start[3], count[3], stride[3];
} kapsel_h5_crop;

/* hdf5_writer::hdf5_writer (output_writer.cxx:41-209): creates the file,
 * its root attributes and system data.  print_list: the nump_print particle
 * ids to dump (NULL: the first nump_print); spec: Particle.spec of all
 * nump_total particles (NULL: 0); crop NULL: the whole grid; with_fields 0:
 * no field group (print_field.none). */
int kapsel_h5_create(const char *out_dir, const char *out_name, int nx, int ny, int nz, double dx, double dt,
                     int nump_total, int nump_print, const int *print_list, const int *spec,
                     const kapsel_h5_crop *crop, int with_fields, kapsel_h5 **out);
/* write_start (output_writer.cxx:242-276): opens frame_<n>, n = frames so far */
int kapsel_h5_frame_begin(kapsel_h5 *h);
/* write_field_scalar (output_writer.cxx:303-307): one padded field under field/<name> */
int kapsel_h5_write_field(kapsel_h5 *h, const char *name, const double *field);
/* write_particle_data (output_writer.cxx:423-520): the particle datasets of
 * the frame; force_h and torque_h are f_hydro_previous / torque_hydro_previous
 * (this library keeps no slip forces); orientations may be NULL (identity) */
int kapsel_h5_write_particles(kapsel_h5 *h, const double *state, const double *orientations);
/* write_end (output_writer.cxx:277-301) */
int kapsel_h5_frame_end(kapsel_h5 *h);
/* ~hdf5_writer (output_writer.cxx:211-240): rewrites the .conf file, closes */
int kapsel_h5_close(kapsel_h5 *h);
/* Reads a dataset ("/trajectory_data/frame_0/field/u_x") or an attribute
 * ("/@dx", "/trajectory_data/frame_0@time") as doubles: *count receives the
 * number of values; out may be NULL to ask for the count only. */
int kapsel_h5_read(const char *path, const char *object, double *out, long long capacity, long long *count);
const char *kapsel_h5_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* KAPSEL_H5_H */
