/* NetCDF-4 files for the optics tables, written and read through HDF5.
 *
 * runoptics writes its tables with netCDF4-python in the NetCDF-4 format
 * (src/geosmie/dointegration.py:205-354: netCDF4.Dataset(..., 'w'),
 * createDimension, createVariable(..., compression='zlib')), the GSF step adds
 * pmom the same way (src/gsf/convertncdf.py:70-79) and the band averaging
 * writes its output likewise (src/geosmie/bandaverage.py:222-290).  NetCDF-4
 * is HDF5 with a fixed set of conventions; this file implements the part of
 * them those writers use, on the HDF5 C library alone:
 *
 *   - one dataset per variable in the root group, link and attribute creation
 *     order tracked and indexed (variables are listed in definition order);
 *   - a variable named after its only dimension (a coordinate variable) is
 *     that dimension's dimension scale (H5DSset_scale with the variable name);
 *     a dimension without one gets an empty 32-bit float dataset of its
 *     length named after it, whose scale NAME is "This is a netCDF dimension
 *     but not a netCDF variable." followed by the length in %10d;
 *   - every scale carries the dimension id in the int attribute _Netcdf4Dimid;
 *     every other variable has its scales attached (DIMENSION_LIST);
 *   - NC_CHAR text attributes are fixed-length, null-terminated C strings in
 *     a scalar dataspace (a null dataspace when empty);
 *   - compressed variables are chunked with the shuffle filter followed by
 *     deflate (netCDF4-python's compression='zlib': shuffle on, level 4);
 *   - the netCDF default fill value of the type is the dataset fill value;
 *     object modification times are not stored.
 *
 * As in the netCDF C API, definitions (dimensions, variables, attributes) come
 * first; the first data write (or gmnc_enddef) creates the datasets.  The
 * reader lists the variables of any file that follows these conventions,
 * with their dimension names, types, text attributes and data.
 *
 * Declarations: include/geosmie_nc4.h.
 */
#include <hdf5.h>
#include <hdf5_hl.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "geosmie_nc4.h"

#define GMNC_MAX_NAME 256
#define GMNC_CHUNK_BYTES (4u << 20)
static const char kPhonyPrefix[] = "This is a netCDF dimension but not a netCDF variable.";

static __thread char g_err[512];

static int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return GMNC_ERR;
}

const char* gmnc_last_error(void) { return g_err; }

typedef struct {
  char* name;
  char* text;
} att_t;

typedef struct {
  char name[GMNC_MAX_NAME];
  int64_t len;
  int coord_var; /* variable id of its coordinate variable, -1 if none */
  hid_t scale;   /* dataset that is its dimension scale (after enddef) */
} dim_t;

typedef struct {
  char name[GMNC_MAX_NAME];
  int type;
  int ndims;
  int dimids[GMNC_MAX_DIMS];
  int deflate, shuffle;
  att_t* atts;
  int n_atts;
  hid_t did;
  /* reader */
  int64_t shape[GMNC_MAX_DIMS];
  char dimnames[GMNC_MAX_DIMS][GMNC_MAX_NAME];
} var_t;

struct gmnc_file {
  hid_t fid;
  int writing;
  int defined; /* datasets created */
  dim_t* dims;
  int n_dims;
  var_t* vars;
  int n_vars;
  att_t* gatts;
  int n_gatts;
};

static void quiet(void) {
  static int done = 0;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    done = 1;
  }
}

static hid_t file_type(int type) {
  switch (type) {
    case GMNC_BYTE: return H5T_STD_I8LE;
    case GMNC_SHORT: return H5T_STD_I16LE;
    case GMNC_INT: return H5T_STD_I32LE;
    case GMNC_FLOAT: return H5T_IEEE_F32LE;
    case GMNC_DOUBLE: return H5T_IEEE_F64LE;
    case GMNC_INT64: return H5T_STD_I64LE;
    default: return -1;
  }
}

static hid_t mem_type(int type) {
  switch (type) {
    case GMNC_BYTE: return H5T_NATIVE_SCHAR;
    case GMNC_SHORT: return H5T_NATIVE_SHORT;
    case GMNC_INT: return H5T_NATIVE_INT;
    case GMNC_FLOAT: return H5T_NATIVE_FLOAT;
    case GMNC_DOUBLE: return H5T_NATIVE_DOUBLE;
    case GMNC_INT64: return H5T_NATIVE_LLONG;
    default: return -1;
  }
}

static size_t type_size(int type) {
  switch (type) {
    case GMNC_BYTE: case GMNC_CHAR: return 1;
    case GMNC_SHORT: return 2;
    case GMNC_INT: case GMNC_FLOAT: return 4;
    case GMNC_DOUBLE: case GMNC_INT64: return 8;
    default: return 0;
  }
}

/* fixed-length, null-terminated C string type of n bytes (n >= 1) */
static hid_t char_type(size_t n) {
  hid_t t = H5Tcopy(H5T_C_S1);
  if (t < 0) return t;
  H5Tset_strpad(t, H5T_STR_NULLTERM);
  H5Tset_size(t, n);
  return t;
}

static int check_name(const char* name) {
  if (!name || !name[0]) return fail("empty name");
  if (strlen(name) >= GMNC_MAX_NAME) return fail("name longer than %d bytes: %.40s...",
                                                 GMNC_MAX_NAME - 1, name);
  if (strchr(name, '/')) return fail("'/' in name %s", name);
  return GMNC_OK;
}

static int add_att(att_t** atts, int* n, const char* name, const char* text) {
  for (int i = 0; i < *n; ++i)
    if (!strcmp((*atts)[i].name, name)) {
      char* t = strdup(text);
      if (!t) return fail("out of memory");
      free((*atts)[i].text);
      (*atts)[i].text = t;
      return GMNC_OK;
    }
  att_t* a = (att_t*)realloc(*atts, sizeof(att_t) * (size_t)(*n + 1));
  if (!a) return fail("out of memory");
  *atts = a;
  a[*n].name = strdup(name);
  a[*n].text = strdup(text);
  if (!a[*n].name || !a[*n].text) return fail("out of memory");
  ++*n;
  return GMNC_OK;
}

static void free_atts(att_t* a, int n) {
  for (int i = 0; i < n; ++i) {
    free(a[i].name);
    free(a[i].text);
  }
  free(a);
}

/* ----------------------------------------------------------------- writer */

int gmnc_create(const char* path, gmnc_file** out) {
  quiet();
  if (!path || !out) return fail("gmnc_create: null argument");
  *out = NULL;
  hid_t fcpl = H5Pcreate(H5P_FILE_CREATE);
  if (fcpl < 0) return fail("gmnc_create: H5Pcreate failed");
  H5Pset_link_creation_order(fcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  H5Pset_attr_creation_order(fcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  hid_t fid = H5Fcreate(path, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT);
  H5Pclose(fcpl);
  if (fid < 0) return fail("gmnc_create: cannot create %s", path);
  gmnc_file* f = (gmnc_file*)calloc(1, sizeof *f);
  if (!f) {
    H5Fclose(fid);
    return fail("out of memory");
  }
  f->fid = fid;
  f->writing = 1;
  *out = f;
  return GMNC_OK;
}

int gmnc_def_dim(gmnc_file* f, const char* name, int64_t len, int* dimid) {
  if (!f || !f->writing || f->defined) return fail("gmnc_def_dim: not in define mode");
  if (check_name(name)) return GMNC_ERR;
  if (len < 0) return fail("gmnc_def_dim: negative length for %s", name);
  for (int i = 0; i < f->n_dims; ++i)
    if (!strcmp(f->dims[i].name, name)) return fail("gmnc_def_dim: %s defined twice", name);
  dim_t* d = (dim_t*)realloc(f->dims, sizeof(dim_t) * (size_t)(f->n_dims + 1));
  if (!d) return fail("out of memory");
  f->dims = d;
  dim_t* x = &d[f->n_dims];
  memset(x, 0, sizeof *x);
  strcpy(x->name, name);
  x->len = len;
  x->coord_var = -1;
  x->scale = -1;
  if (dimid) *dimid = f->n_dims;
  ++f->n_dims;
  return GMNC_OK;
}

int gmnc_def_var(gmnc_file* f, const char* name, int type, int ndims, const int* dimids,
                 int deflate_level, int shuffle, int* varid) {
  if (!f || !f->writing || f->defined) return fail("gmnc_def_var: not in define mode");
  if (check_name(name)) return GMNC_ERR;
  if (type != GMNC_CHAR && file_type(type) < 0) return fail("gmnc_def_var: bad type %d", type);
  if (ndims < 0 || ndims > GMNC_MAX_DIMS) return fail("gmnc_def_var: %d dimensions", ndims);
  if (deflate_level < 0 || deflate_level > 9) return fail("gmnc_def_var: deflate level %d",
                                                          deflate_level);
  for (int i = 0; i < f->n_vars; ++i)
    if (!strcmp(f->vars[i].name, name)) return fail("gmnc_def_var: %s defined twice", name);
  for (int i = 0; i < ndims; ++i)
    if (!dimids || dimids[i] < 0 || dimids[i] >= f->n_dims)
      return fail("gmnc_def_var: %s: bad dimension id", name);
  var_t* v = (var_t*)realloc(f->vars, sizeof(var_t) * (size_t)(f->n_vars + 1));
  if (!v) return fail("out of memory");
  f->vars = v;
  var_t* x = &v[f->n_vars];
  memset(x, 0, sizeof *x);
  strcpy(x->name, name);
  x->type = type;
  x->ndims = ndims;
  for (int i = 0; i < ndims; ++i) x->dimids[i] = dimids[i];
  x->deflate = deflate_level;
  x->shuffle = shuffle ? 1 : 0;
  x->did = -1;
  if (ndims == 1 && !strcmp(f->dims[dimids[0]].name, name)) f->dims[dimids[0]].coord_var = f->n_vars;
  if (varid) *varid = f->n_vars;
  ++f->n_vars;
  return GMNC_OK;
}

int gmnc_put_att_text(gmnc_file* f, int varid, const char* name, const char* text) {
  if (!f || !f->writing || f->defined) return fail("gmnc_put_att_text: not in define mode");
  if (check_name(name)) return GMNC_ERR;
  if (!text) return fail("gmnc_put_att_text: null text");
  if (varid == GMNC_GLOBAL) return add_att(&f->gatts, &f->n_gatts, name, text);
  if (varid < 0 || varid >= f->n_vars) return fail("gmnc_put_att_text: bad variable id");
  return add_att(&f->vars[varid].atts, &f->vars[varid].n_atts, name, text);
}

static int write_text_att(hid_t obj, const char* name, const char* text) {
  size_t n = strlen(text);
  hid_t t = char_type(n ? n : 1);
  hid_t s = n ? H5Screate(H5S_SCALAR) : H5Screate(H5S_NULL);
  hid_t a = H5Acreate2(obj, name, t, s, H5P_DEFAULT, H5P_DEFAULT);
  int rc = GMNC_OK;
  if (a < 0 || (n && H5Awrite(a, t, text) < 0)) rc = fail("attribute %s not written", name);
  if (a >= 0) H5Aclose(a);
  H5Sclose(s);
  H5Tclose(t);
  return rc;
}

static int write_int_att(hid_t obj, const char* name, int value) {
  hid_t s = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(obj, name, H5T_NATIVE_INT, s, H5P_DEFAULT, H5P_DEFAULT);
  int rc = GMNC_OK;
  if (a < 0 || H5Awrite(a, H5T_NATIVE_INT, &value) < 0) rc = fail("attribute %s not written", name);
  if (a >= 0) H5Aclose(a);
  H5Sclose(s);
  return rc;
}

/* chunk shape: the whole variable, leading dimensions halved until one chunk
 * holds at most 4 MiB; every extent at least 1 */
static void chunk_shape(const hsize_t* dims, int nd, size_t esz, hsize_t* chunk) {
  uint64_t bytes = esz;
  for (int i = 0; i < nd; ++i) {
    chunk[i] = dims[i] ? dims[i] : 1;
    bytes *= chunk[i];
  }
  for (int i = 0; i < nd && bytes > GMNC_CHUNK_BYTES; ++i) {
    while (chunk[i] > 1 && bytes > GMNC_CHUNK_BYTES) {
      uint64_t rest = bytes / chunk[i];
      chunk[i] = (chunk[i] + 1) / 2;
      bytes = rest * chunk[i];
    }
  }
}

static int create_var(gmnc_file* f, int vi) {
  var_t* v = &f->vars[vi];
  hsize_t dims[GMNC_MAX_DIMS], chunk[GMNC_MAX_DIMS];
  for (int i = 0; i < v->ndims; ++i) dims[i] = (hsize_t)f->dims[v->dimids[i]].len;
  hid_t space = v->ndims ? H5Screate_simple(v->ndims, dims, dims) : H5Screate(H5S_SCALAR);
  hid_t ft = v->type == GMNC_CHAR ? char_type(1) : H5Tcopy(file_type(v->type));
  hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
  H5Pset_attr_creation_order(dcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  H5Pset_obj_track_times(dcpl, 0);
  if (v->deflate > 0 && v->ndims > 0) {
    chunk_shape(dims, v->ndims, type_size(v->type), chunk);
    H5Pset_chunk(dcpl, v->ndims, chunk);
    if (v->shuffle) H5Pset_shuffle(dcpl);
    H5Pset_deflate(dcpl, (unsigned)v->deflate);
  }
  /* netCDF default fill values (NC_FILL_*) */
  switch (v->type) {
    case GMNC_BYTE: { signed char x = -127; H5Pset_fill_value(dcpl, H5T_NATIVE_SCHAR, &x); break; }
    case GMNC_CHAR: { char x = 0; H5Pset_fill_value(dcpl, ft, &x); break; }
    case GMNC_SHORT: { short x = -32767; H5Pset_fill_value(dcpl, H5T_NATIVE_SHORT, &x); break; }
    case GMNC_INT: { int x = -2147483647; H5Pset_fill_value(dcpl, H5T_NATIVE_INT, &x); break; }
    case GMNC_FLOAT: { float x = 9.9692099683868690e+36f; H5Pset_fill_value(dcpl, H5T_NATIVE_FLOAT, &x); break; }
    case GMNC_DOUBLE: { double x = 9.9692099683868690e+36; H5Pset_fill_value(dcpl, H5T_NATIVE_DOUBLE, &x); break; }
    case GMNC_INT64: { long long x = -9223372036854775806LL; H5Pset_fill_value(dcpl, H5T_NATIVE_LLONG, &x); break; }
  }
  v->did = H5Dcreate2(f->fid, v->name, ft, space, H5P_DEFAULT, dcpl, H5P_DEFAULT);
  H5Pclose(dcpl);
  H5Tclose(ft);
  H5Sclose(space);
  if (v->did < 0) return fail("dataset %s not created", v->name);
  for (int i = 0; i < v->n_atts; ++i)
    if (write_text_att(v->did, v->atts[i].name, v->atts[i].text)) return GMNC_ERR;
  return GMNC_OK;
}

static int create_phony_dim(gmnc_file* f, int di) {
  dim_t* d = &f->dims[di];
  for (int i = 0; i < f->n_vars; ++i)
    if (!strcmp(f->vars[i].name, d->name))
      return fail("variable %s has the name of dimension %s but is not its coordinate variable",
                  d->name, d->name);
  hsize_t n = (hsize_t)d->len;
  hid_t space = H5Screate_simple(1, &n, &n);
  hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
  H5Pset_attr_creation_order(dcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  H5Pset_obj_track_times(dcpl, 0);
  d->scale = H5Dcreate2(f->fid, d->name, H5T_IEEE_F32BE, space, H5P_DEFAULT, dcpl, H5P_DEFAULT);
  H5Pclose(dcpl);
  H5Sclose(space);
  if (d->scale < 0) return fail("dimension %s not created", d->name);
  char nm[sizeof kPhonyPrefix + 16];
  snprintf(nm, sizeof nm, "%s%10d", kPhonyPrefix, (int)d->len);
  if (H5DSset_scale(d->scale, nm) < 0) return fail("dimension scale %s not set", d->name);
  return GMNC_OK;
}

int gmnc_enddef(gmnc_file* f) {
  if (!f || !f->writing) return fail("gmnc_enddef: file not open for writing");
  if (f->defined) return GMNC_OK;
  f->defined = 1;
  for (int i = 0; i < f->n_gatts; ++i)
    if (write_text_att(f->fid, f->gatts[i].name, f->gatts[i].text)) return GMNC_ERR;
  for (int vi = 0; vi < f->n_vars; ++vi)
    if (create_var(f, vi)) return GMNC_ERR;
  for (int di = 0; di < f->n_dims; ++di) {
    dim_t* d = &f->dims[di];
    if (d->coord_var >= 0) {
      d->scale = f->vars[d->coord_var].did;
      if (H5DSset_scale(d->scale, d->name) < 0) return fail("dimension scale %s not set", d->name);
    } else if (create_phony_dim(f, di)) {
      return GMNC_ERR;
    }
    if (write_int_att(d->scale, "_Netcdf4Dimid", di)) return GMNC_ERR;
  }
  for (int vi = 0; vi < f->n_vars; ++vi) {
    var_t* v = &f->vars[vi];
    if (v->ndims == 1 && f->dims[v->dimids[0]].coord_var == vi) continue;
    for (int i = 0; i < v->ndims; ++i)
      if (H5DSattach_scale(v->did, f->dims[v->dimids[i]].scale, (unsigned)i) < 0)
        return fail("dimension %s not attached to %s", f->dims[v->dimids[i]].name, v->name);
  }
  return GMNC_OK;
}

int gmnc_put_var(gmnc_file* f, int varid, const void* data) {
  if (!f || !f->writing) return fail("gmnc_put_var: file not open for writing");
  if (varid < 0 || varid >= f->n_vars) return fail("gmnc_put_var: bad variable id");
  if (!f->defined && gmnc_enddef(f)) return GMNC_ERR;
  var_t* v = &f->vars[varid];
  int64_t n = 1;
  for (int i = 0; i < v->ndims; ++i) n *= f->dims[v->dimids[i]].len;
  if (n == 0) return GMNC_OK;
  if (!data) return fail("gmnc_put_var: null data for %s", v->name);
  hid_t mt = v->type == GMNC_CHAR ? char_type(1) : H5Tcopy(mem_type(v->type));
  herr_t e = H5Dwrite(v->did, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  H5Tclose(mt);
  if (e < 0) return fail("data of %s not written", v->name);
  return GMNC_OK;
}

/* ----------------------------------------------------------------- reader */

typedef struct {
  gmnc_file* f;
  int rc;
} iter_t;

static int read_text_att(hid_t obj, const char* name, char* buf, size_t cap, size_t* len) {
  hid_t a = H5Aopen(obj, name, H5P_DEFAULT);
  if (a < 0) return fail("attribute %s missing", name);
  hid_t t = H5Aget_type(a), s = H5Aget_space(a);
  int rc = GMNC_OK;
  if (H5Tget_class(t) != H5T_STRING) {
    rc = fail("attribute %s is not text", name);
  } else if (H5Sget_simple_extent_type(s) == H5S_NULL) {
    *len = 0;
    if (cap) buf[0] = 0;
  } else if (H5Sget_simple_extent_npoints(s) != 1) {
    rc = fail("attribute %s: text arrays are not supported", name);
  } else if (H5Tis_variable_str(t) > 0) {
    hid_t mt = H5Tcopy(H5T_C_S1);
    H5Tset_size(mt, H5T_VARIABLE);
    char* p = NULL;
    if (H5Aread(a, mt, &p) < 0) {
      rc = fail("attribute %s not read", name);
    } else {
      *len = p ? strlen(p) : 0;
      if (cap) {
        size_t k = *len < cap - 1 ? *len : cap - 1;
        if (k) memcpy(buf, p, k);
        buf[k] = 0;
      }
      H5free_memory(p);
    }
    H5Tclose(mt);
  } else {
    size_t n = H5Tget_size(t);
    char* tmp = (char*)calloc(n + 1, 1);
    hid_t mt = H5Tcopy(t);
    if (!tmp || H5Aread(a, mt, tmp) < 0) {
      rc = fail("attribute %s not read", name);
    } else {
      *len = strnlen(tmp, n);
      if (cap) {
        size_t k = *len < cap - 1 ? *len : cap - 1;
        memcpy(buf, tmp, k);
        buf[k] = 0;
      }
    }
    free(tmp);
    H5Tclose(mt);
  }
  H5Sclose(s);
  H5Tclose(t);
  H5Aclose(a);
  return rc;
}

static int is_phony(hid_t did) {
  if (H5Aexists(did, "CLASS") <= 0 || H5Aexists(did, "NAME") <= 0) return 0;
  char nm[GMNC_MAX_NAME];
  size_t len = 0;
  if (read_text_att(did, "NAME", nm, sizeof nm, &len)) return 0;
  return !strncmp(nm, kPhonyPrefix, sizeof kPhonyPrefix - 1);
}

static int classify(hid_t did, int* type) {
  hid_t t = H5Dget_type(did);
  H5T_class_t c = H5Tget_class(t);
  size_t n = H5Tget_size(t);
  int rc = GMNC_OK;
  if (c == H5T_FLOAT && n == 8) *type = GMNC_DOUBLE;
  else if (c == H5T_FLOAT && n == 4) *type = GMNC_FLOAT;
  else if (c == H5T_INTEGER && H5Tget_sign(t) == H5T_SGN_2) {
    if (n == 8) *type = GMNC_INT64;
    else if (n == 4) *type = GMNC_INT;
    else if (n == 2) *type = GMNC_SHORT;
    else if (n == 1) *type = GMNC_BYTE;
    else rc = GMNC_ERR;
  } else if (c == H5T_STRING && n == 1 && H5Tis_variable_str(t) <= 0) *type = GMNC_CHAR;
  else rc = GMNC_ERR;
  H5Tclose(t);
  return rc;
}

static herr_t scale_name_cb(hid_t did, unsigned dim, hid_t scale, void* data) {
  (void)did;
  (void)dim;
  char* out = (char*)data;
  char path[GMNC_MAX_NAME + 2];
  ssize_t n = H5Iget_name(scale, path, sizeof path);
  if (n <= 0) return -1;
  const char* p = strrchr(path, '/');
  p = p ? p + 1 : path;
  strncpy(out, p, GMNC_MAX_NAME - 1);
  out[GMNC_MAX_NAME - 1] = 0;
  return 1;
}

static herr_t list_cb(hid_t gid, const char* name, const H5L_info_t* info, void* data) {
  iter_t* it = (iter_t*)data;
  gmnc_file* f = it->f;
  if (info->type != H5L_TYPE_HARD) return 0;
  hid_t oid = H5Oopen(gid, name, H5P_DEFAULT);
  if (oid < 0) return 0;
  if (H5Iget_type(oid) != H5I_DATASET || is_phony(oid)) {
    H5Oclose(oid);
    return 0;
  }
  int type;
  if (classify(oid, &type)) {  /* a type the tables never use: not listed */
    H5Oclose(oid);
    return 0;
  }
  if (strlen(name) >= GMNC_MAX_NAME) {
    H5Oclose(oid);
    it->rc = fail("variable name too long: %.40s...", name);
    return -1;
  }
  var_t* v = (var_t*)realloc(f->vars, sizeof(var_t) * (size_t)(f->n_vars + 1));
  if (!v) {
    H5Oclose(oid);
    it->rc = fail("out of memory");
    return -1;
  }
  f->vars = v;
  var_t* x = &v[f->n_vars++];
  memset(x, 0, sizeof *x);
  strcpy(x->name, name);
  x->type = type;
  x->did = oid;
  hid_t space = H5Dget_space(oid);
  int nd = H5Sget_simple_extent_ndims(space);
  if (nd < 0 || nd > GMNC_MAX_DIMS) {
    H5Sclose(space);
    it->rc = fail("variable %s: %d dimensions", name, nd);
    return -1;
  }
  hsize_t dims[GMNC_MAX_DIMS];
  H5Sget_simple_extent_dims(space, dims, NULL);
  H5Sclose(space);
  x->ndims = nd;
  const int self_scale = H5DSis_scale(oid) > 0;
  for (int i = 0; i < nd; ++i) {
    x->shape[i] = (int64_t)dims[i];
    x->dimnames[i][0] = 0;
    if (self_scale && nd == 1) {
      strcpy(x->dimnames[i], name);
    } else if (!self_scale && H5DSget_num_scales(oid, (unsigned)i) > 0) {
      H5DSiterate_scales(oid, (unsigned)i, NULL, scale_name_cb, x->dimnames[i]);
    }
  }
  return 0;
}

int gmnc_open(const char* path, gmnc_file** out) {
  quiet();
  if (!path || !out) return fail("gmnc_open: null argument");
  *out = NULL;
  hid_t fid = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (fid < 0) return fail("gmnc_open: %s is not an HDF5 (NetCDF-4) file", path);
  gmnc_file* f = (gmnc_file*)calloc(1, sizeof *f);
  if (!f) {
    H5Fclose(fid);
    return fail("out of memory");
  }
  f->fid = fid;
  iter_t it = {f, GMNC_OK};
  hsize_t idx = 0;
  herr_t e = H5Literate(fid, H5_INDEX_CRT_ORDER, H5_ITER_INC, &idx, list_cb, &it);
  if (e < 0 && it.rc == GMNC_OK) {
    /* no creation-order index in this file: list by name */
    for (int i = 0; i < f->n_vars; ++i) H5Dclose(f->vars[i].did);
    free(f->vars);
    f->vars = NULL;
    f->n_vars = 0;
    idx = 0;
    e = H5Literate(fid, H5_INDEX_NAME, H5_ITER_INC, &idx, list_cb, &it);
  }
  if (e < 0 || it.rc != GMNC_OK) {
    gmnc_close(f);
    if (it.rc == GMNC_OK) fail("gmnc_open: cannot list the variables of %s", path);
    return GMNC_ERR;
  }
  *out = f;
  return GMNC_OK;
}

int gmnc_inq_nvars(const gmnc_file* f, int* nvars) {
  if (!f || !nvars) return fail("gmnc_inq_nvars: null argument");
  *nvars = f->n_vars;
  return GMNC_OK;
}

static const var_t* rvar(const gmnc_file* f, int varid) {
  if (!f || varid < 0 || varid >= f->n_vars) {
    fail("bad variable id %d", varid);
    return NULL;
  }
  return &f->vars[varid];
}

int gmnc_inq_var(const gmnc_file* f, int varid, char* name, size_t name_cap, int* type,
                 int* ndims, int64_t* shape) {
  const var_t* v = rvar(f, varid);
  if (!v) return GMNC_ERR;
  if (name && name_cap) {
    strncpy(name, v->name, name_cap - 1);
    name[name_cap - 1] = 0;
  }
  if (type) *type = v->type;
  if (ndims) *ndims = v->ndims;
  if (shape)
    for (int i = 0; i < v->ndims; ++i) shape[i] = f->writing ? f->dims[v->dimids[i]].len
                                                             : v->shape[i];
  return GMNC_OK;
}

int gmnc_inq_dimname(const gmnc_file* f, int varid, int dim, char* name, size_t name_cap) {
  const var_t* v = rvar(f, varid);
  if (!v) return GMNC_ERR;
  if (dim < 0 || dim >= v->ndims || !name || !name_cap) return fail("gmnc_inq_dimname: bad dim");
  const char* s = f->writing ? f->dims[v->dimids[dim]].name : v->dimnames[dim];
  strncpy(name, s, name_cap - 1);
  name[name_cap - 1] = 0;
  return GMNC_OK;
}

int gmnc_inq_filters(const gmnc_file* f, int varid, int* shuffle, int* deflate_level) {
  const var_t* v = rvar(f, varid);
  if (!v) return GMNC_ERR;
  if (v->did < 0) return fail("gmnc_inq_filters: datasets not created yet");
  *shuffle = 0;
  *deflate_level = 0;
  hid_t dcpl = H5Dget_create_plist(v->did);
  int n = H5Pget_nfilters(dcpl);
  for (int i = 0; i < n; ++i) {
    unsigned flags, cd[8];
    size_t ncd = 8;
    char fname[64];
    unsigned fc;
    H5Z_filter_t id = H5Pget_filter2(dcpl, (unsigned)i, &flags, &ncd, cd, sizeof fname, fname, &fc);
    if (id == H5Z_FILTER_SHUFFLE) *shuffle = 1;
    if (id == H5Z_FILTER_DEFLATE) *deflate_level = ncd ? (int)cd[0] : -1;
  }
  H5Pclose(dcpl);
  return GMNC_OK;
}

int gmnc_get_att_text(const gmnc_file* f, int varid, const char* name, char* buf, size_t cap,
                      size_t* len) {
  if (!f || !name || !len) return fail("gmnc_get_att_text: null argument");
  hid_t obj;
  if (varid == GMNC_GLOBAL) {
    obj = f->fid;
  } else {
    const var_t* v = rvar(f, varid);
    if (!v) return GMNC_ERR;
    if (v->did < 0) return fail("gmnc_get_att_text: datasets not created yet");
    obj = v->did;
  }
  if (H5Aexists(obj, name) <= 0) return GMNC_NOTFOUND;
  return read_text_att(obj, name, buf, cap, len);
}

int gmnc_get_var(const gmnc_file* f, int varid, void* data) {
  const var_t* v = rvar(f, varid);
  if (!v) return GMNC_ERR;
  if (v->did < 0) return fail("gmnc_get_var: datasets not created yet");
  int64_t n = 1;
  for (int i = 0; i < v->ndims; ++i) n *= f->writing ? f->dims[v->dimids[i]].len : v->shape[i];
  if (n == 0) return GMNC_OK;
  if (!data) return fail("gmnc_get_var: null buffer for %s", v->name);
  /* characters: the file's own string type, so the bytes pass unconverted
   * whatever its padding (a null-terminated 1-byte string would read as 0) */
  hid_t mt = v->type == GMNC_CHAR ? H5Dget_type(v->did) : H5Tcopy(mem_type(v->type));
  herr_t e = H5Dread(v->did, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  H5Tclose(mt);
  if (e < 0) return fail("data of %s not read", v->name);
  return GMNC_OK;
}

int gmnc_close(gmnc_file* f) {
  if (!f) return GMNC_OK;
  int rc = GMNC_OK;
  if (f->writing && !f->defined) rc = gmnc_enddef(f);
  for (int i = 0; i < f->n_dims; ++i)
    if (f->dims[i].coord_var < 0 && f->dims[i].scale >= 0) H5Dclose(f->dims[i].scale);
  for (int i = 0; i < f->n_vars; ++i) {
    if (f->vars[i].did >= 0) H5Dclose(f->vars[i].did);
    free_atts(f->vars[i].atts, f->vars[i].n_atts);
  }
  free_atts(f->gatts, f->n_gatts);
  if (f->fid >= 0 && H5Fclose(f->fid) < 0 && rc == GMNC_OK) rc = fail("file not closed cleanly");
  free(f->dims);
  free(f->vars);
  free(f);
  return rc;
}
