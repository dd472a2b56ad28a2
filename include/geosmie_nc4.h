/* NetCDF-4 table files: a writer and reader on the HDF5 C library
 * (libgeosmie_nc4.so, plain C, no GPU).
 *
 * The reference writes its optics tables, the GSF moments and the band
 * averages with netCDF4-python in the NetCDF-4 format:
 *   src/geosmie/dointegration.py:205-354  createNCDF (dimensions rh, wavelength,
 *                                         bin, p, ang; i8 bin and p indices;
 *                                         f8 variables, compression='zlib')
 *   src/geosmie/dointegration.py:1006-1031 the per-entry writes
 *   src/gsf/convertncdf.py:70-79           pmom on a new dimension m / nMom
 *   src/geosmie/bandaverage.py:222-290     the band-averaged file
 * These functions replace the netCDF4.Dataset calls those lines make
 * (createDimension, createVariable, the attribute and slice assignments,
 * close; Dataset(path, 'r') and variables[...][:] for reading), with the
 * netCDF C API's define-then-write model and nc_type codes.
 *
 * Return values: GMNC_OK (0), GMNC_ERR (-1, message in gmnc_last_error()),
 * GMNC_NOTFOUND (1, gmnc_get_att_text only).
 */
#ifndef GEOSMIE_NC4_H
#define GEOSMIE_NC4_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMNC_OK 0
#define GMNC_ERR (-1)
#define GMNC_NOTFOUND 1

#define GMNC_GLOBAL (-1) /* variable id of the file's own attributes */
#define GMNC_MAX_DIMS 8

/* netCDF's nc_type codes */
enum gmnc_type {
  GMNC_BYTE = 1,
  GMNC_CHAR = 2,
  GMNC_SHORT = 3,
  GMNC_INT = 4,
  GMNC_FLOAT = 5,
  GMNC_DOUBLE = 6,
  GMNC_INT64 = 10
};

typedef struct gmnc_file gmnc_file;

const char* gmnc_last_error(void);

/* netCDF4.Dataset(path, 'w'): create (truncate) a NetCDF-4 file */
int gmnc_create(const char* path, gmnc_file** f);
/* createDimension(name, len) */
int gmnc_def_dim(gmnc_file* f, const char* name, int64_t len, int* dimid);
/* createVariable(name, type, dims, compression='zlib' when deflate_level > 0;
 * shuffle as netCDF4-python's shuffle=True) */
int gmnc_def_var(gmnc_file* f, const char* name, int type, int ndims, const int* dimids,
                 int deflate_level, int shuffle, int* varid);
/* variables[name].units = text (varid GMNC_GLOBAL: a file attribute) */
int gmnc_put_att_text(gmnc_file* f, int varid, const char* name, const char* text);
/* end of definitions: creates every dataset (called by the first put) */
int gmnc_enddef(gmnc_file* f);
/* variables[name][:] = data (C order, the variable's type in native layout) */
int gmnc_put_var(gmnc_file* f, int varid, const void* data);

/* netCDF4.Dataset(path, 'r'): the variables in definition order */
int gmnc_open(const char* path, gmnc_file** f);
int gmnc_inq_nvars(const gmnc_file* f, int* nvars);
/* name, type, number of dimensions and their lengths (shape: GMNC_MAX_DIMS) */
int gmnc_inq_var(const gmnc_file* f, int varid, char* name, size_t name_cap, int* type,
                 int* ndims, int64_t* shape);
int gmnc_inq_dimname(const gmnc_file* f, int varid, int dim, char* name, size_t name_cap);
/* shuffle on/off and deflate level (0: not compressed) */
int gmnc_inq_filters(const gmnc_file* f, int varid, int* shuffle, int* deflate_level);
/* text attribute; *len = its length (the copy in buf is cut to cap - 1 bytes) */
int gmnc_get_att_text(const gmnc_file* f, int varid, const char* name, char* buf, size_t cap,
                      size_t* len);
int gmnc_get_var(const gmnc_file* f, int varid, void* data);

/* Dataset.close(); writes pending definitions first */
int gmnc_close(gmnc_file* f);

#ifdef __cplusplus
}
#endif

#endif
