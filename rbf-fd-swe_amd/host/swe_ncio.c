/*
 * NetCDF input of the RBF-FD SWE host (the reference's USE_NCIO path,
 * swe_code/io/input.c:14-186) and the nc -> bin converter
 * (swe_code/io/nc2bin.c:10-84), without libnetcdf: a reader of the NetCDF
 * classic file formats written from their published specification
 * ("NetCDF Classic and 64-bit Offset File Formats", Unidata; CDF-5 from the
 * PnetCDF/Unidata CDF-5 format note):
 *
 *   magic 'C' 'D' 'F' v   v = 1 classic, 2 64-bit offset, 5 64-bit data
 *   numrecs                  int32 (v 1, 2) or int64 (v 5), big-endian
 *   dim_list, gatt_list, var_list: ABSENT (two zero words) or
 *       tag (0x0A dims, 0x0C atts, 0x0B vars) + count + entries
 *   dim  = name + length (0 = the record dimension)
 *   attr = name + type + count + values padded to 4 bytes
 *   var  = name + ndims + dimids + attrs + type + vsize + begin
 *          (begin: int32 in v 1, int64 in v 2 and 5)
 *   data big-endian, C order; a record variable's record r lies at
 *   begin + r * (sum of the record variables' vsize).
 *
 * NetCDF-4 files are HDF5 containers and are rejected with
 * SWE_HOST_EFORMAT (no HDF5 library in this build).
 *
 * Variables read, as the reference names them: Nnodes, Nnbr (int), a, g,
 * gh0, timestep, gamma (scalars), x, y, z, f, ghm [Nnodes], p_u, p_v, p_w,
 * gradghm [Nnodes][3], DPx, DPy, DPz, L, idx [Nnodes][Nnbr], H and H_100
 * [Nnodes][4].  Values of any numeric type are converted as nc_get_var_double
 * and nc_get_var_int convert them (integers out of int range are an error,
 * NC_ERANGE).  Unlike the reference, a variable whose size differs from the
 * size implied by Nnodes and Nnbr is an error (the reference reads it into a
 * buffer of the implied size), and so is a missing idx (the reference does
 * not check that call).
 */
#include "swe_host.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/types.h>

enum { NC_BYTE = 1, NC_CHAR, NC_SHORT, NC_INT, NC_FLOAT, NC_DOUBLE, NC_UBYTE, NC_USHORT,
       NC_UINT, NC_INT64, NC_UINT64 };
enum { TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C };

typedef struct ncvar {
    char *name;
    int type, ndims, isrec;
    long long *dimids;
    long long vsize, begin, nelems;   /* nelems: per record for record variables */
} ncvar;

typedef struct ncfile {
    FILE *fp;
    long long fsize;   /* bytes in the file: bounds every count of the header */
    int version;
    long long numrecs, recsize;
    long long ndims, nvars;
    long long *dimlen;
    ncvar *vars;
} ncfile;

static int type_size(int t)
{
    switch (t) {
    case NC_BYTE: case NC_CHAR: case NC_UBYTE: return 1;
    case NC_SHORT: case NC_USHORT: return 2;
    case NC_INT: case NC_FLOAT: case NC_UINT: return 4;
    case NC_DOUBLE: case NC_INT64: case NC_UINT64: return 8;
    default: return 0;
    }
}

static uint64_t be(const unsigned char *p, int n)
{
    uint64_t v = 0;
    for (int k = 0; k < n; k++)
        v = v << 8 | p[k];
    return v;
}

static int rd(FILE *fp, void *dst, size_t n)
{
    return fread(dst, 1, n, fp) == n ? 0 : SWE_HOST_EREAD;
}

/* a 32-bit big-endian word, or a 64-bit one where the format widens it */
static int rd_count(ncfile *nc, int wide, long long *out)
{
    unsigned char b[8];
    const int n = wide ? 8 : 4;
    if (rd(nc->fp, b, (size_t)n))
        return SWE_HOST_EREAD;
    const uint64_t v = be(b, n);
    if (!wide) {
        *out = (long long)(int32_t)(uint32_t)v;
        return 0;
    }
    if (v > (uint64_t)INT64_MAX)
        return SWE_HOST_EFORMAT;
    *out = (long long)v;
    return 0;
}

static int rd_u32(ncfile *nc, uint32_t *out)
{
    unsigned char b[4];
    if (rd(nc->fp, b, 4))
        return SWE_HOST_EREAD;
    *out = (uint32_t)be(b, 4);
    return 0;
}

static int skip(FILE *fp, long long n)
{
    return fseeko(fp, (off_t)n, SEEK_CUR) == 0 ? 0 : SWE_HOST_EREAD;
}

static long long pad4(long long n) { return (n + 3) & ~3LL; }

/* Byte counts of a file are below 2^62: anything larger is a malformed header. */
#define NC_MAX_BYTES (1LL << 62)

static int rd_name(ncfile *nc, char **out)
{
    long long len;
    int err = rd_count(nc, nc->version == 5, &len);
    if (err)
        return err;
    if (len < 0 || len > 4096)
        return SWE_HOST_EFORMAT;
    char *s = (char *)malloc((size_t)len + 1);
    if (!s)
        return SWE_HOST_ENOMEM;
    if (rd(nc->fp, s, (size_t)len) || skip(nc->fp, pad4(len) - len)) {
        free(s);
        return SWE_HOST_EREAD;
    }
    s[len] = 0;
    *out = s;
    return 0;
}

/* tag + count of a list, or ABSENT (zero tag, zero count) */
static int rd_list_head(ncfile *nc, uint32_t want, long long *count)
{
    uint32_t tag;
    int err = rd_u32(nc, &tag);
    if (!err)
        err = rd_count(nc, nc->version == 5, count);
    if (err)
        return err;
    if (tag == 0 && *count == 0)
        return 0;
    if (tag != want || *count < 0)
        return SWE_HOST_EFORMAT;
    return 0;
}

static int skip_atts(ncfile *nc)
{
    long long n;
    int err = rd_list_head(nc, TAG_ATT, &n);
    for (long long a = 0; a < n && !err; a++) {
        char *name = NULL;
        err = rd_name(nc, &name);
        free(name);
        uint32_t t;
        long long cnt;
        if (!err)
            err = rd_u32(nc, &t);
        if (!err)
            err = rd_count(nc, nc->version == 5, &cnt);
        if (!err && (type_size((int)t) == 0 || cnt < 0 || cnt > NC_MAX_BYTES / 8))
            err = SWE_HOST_EFORMAT;
        if (!err)
            err = skip(nc->fp, pad4(cnt * type_size((int)t)));
    }
    return err;
}

static void nc_close_file(ncfile *nc)
{
    if (nc->fp)
        fclose(nc->fp);
    for (long long v = 0; v < nc->nvars && nc->vars; v++) {
        free(nc->vars[v].name);
        free(nc->vars[v].dimids);
    }
    free(nc->vars);
    free(nc->dimlen);
    memset(nc, 0, sizeof(*nc));
}

static int nc_open_file(const char *path, ncfile *nc)
{
    memset(nc, 0, sizeof(*nc));
    nc->fp = fopen(path, "rb");
    if (!nc->fp)
        return SWE_HOST_EOPEN;
    if (fseeko(nc->fp, 0, SEEK_END) != 0 || (nc->fsize = (long long)ftello(nc->fp)) < 0
        || fseeko(nc->fp, 0, SEEK_SET) != 0) {
        nc_close_file(nc);
        return SWE_HOST_EREAD;
    }
    unsigned char magic[8];
    int err = rd(nc->fp, magic, 4);
    if (!err && memcmp(magic, "\x89HDF", 4) == 0)
        err = SWE_HOST_EFORMAT;   /* NetCDF-4 / HDF5 */
    if (!err && (memcmp(magic, "CDF", 3) != 0
                 || (magic[3] != 1 && magic[3] != 2 && magic[3] != 5)))
        err = SWE_HOST_EFORMAT;
    if (err) {
        nc_close_file(nc);
        return err == SWE_HOST_EREAD ? SWE_HOST_EFORMAT : err;
    }
    nc->version = magic[3];
    const int wide = nc->version == 5;
    uint64_t numrecs;
    if (rd(nc->fp, magic, wide ? 8 : 4)) {
        nc_close_file(nc);
        return SWE_HOST_EFORMAT;
    }
    numrecs = be(magic, wide ? 8 : 4);
    const int streaming = wide ? numrecs == UINT64_MAX : numrecs == 0xFFFFFFFFu;
    if (!streaming && numrecs > (uint64_t)NC_MAX_BYTES) {
        nc_close_file(nc);
        return SWE_HOST_EFORMAT;
    }
    nc->numrecs = streaming ? -1 : (long long)numrecs;
    /* dimensions */
    err = rd_list_head(nc, TAG_DIM, &nc->ndims);
    if (!err && nc->ndims > nc->fsize / 8)   /* a dimension takes at least 8 bytes */
        err = SWE_HOST_EFORMAT;
    if (!err && nc->ndims > 0) {
        nc->dimlen = (long long *)calloc((size_t)nc->ndims, sizeof(long long));
        if (!nc->dimlen)
            err = SWE_HOST_ENOMEM;
    }
    for (long long d = 0; d < nc->ndims && !err; d++) {
        char *name = NULL;
        err = rd_name(nc, &name);
        free(name);
        if (!err)
            err = rd_count(nc, wide, &nc->dimlen[d]);
        if (!err && nc->dimlen[d] < 0)
            err = SWE_HOST_EFORMAT;
    }
    if (!err)
        err = skip_atts(nc);   /* global attributes */
    if (!err)
        err = rd_list_head(nc, TAG_VAR, &nc->nvars);
    if (!err && nc->nvars > nc->fsize / 16)   /* a variable takes at least 16 bytes */
        err = SWE_HOST_EFORMAT;
    if (!err && nc->nvars > 0) {
        nc->vars = (ncvar *)calloc((size_t)nc->nvars, sizeof(ncvar));
        if (!nc->vars)
            err = SWE_HOST_ENOMEM;
    }
    for (long long v = 0; v < nc->nvars && !err; v++) {
        ncvar *x = &nc->vars[v];
        err = rd_name(nc, &x->name);
        long long nd = 0;
        if (!err)
            err = rd_count(nc, wide, &nd);
        if (!err && (nd < 0 || nd > 1024 || nd > nc->fsize / 4))
            err = SWE_HOST_EFORMAT;
        if (!err) {
            x->ndims = (int)nd;
            x->dimids = (long long *)calloc((size_t)(nd ? nd : 1), sizeof(long long));
            if (!x->dimids)
                err = SWE_HOST_ENOMEM;
        }
        x->nelems = 1;
        for (int k = 0; k < x->ndims && !err; k++) {
            err = rd_count(nc, wide, &x->dimids[k]);
            if (!err && (x->dimids[k] < 0 || x->dimids[k] >= nc->ndims))
                err = SWE_HOST_EFORMAT;
            if (err)
                break;
            const long long len = nc->dimlen[x->dimids[k]];
            if (len == 0) {
                if (k != 0)   /* only the first dimension may be the record one */
                    err = SWE_HOST_EFORMAT;
                x->isrec = 1;
            } else if (x->nelems > NC_MAX_BYTES / 8 / len) {
                err = SWE_HOST_EFORMAT;   /* element count past any file size */
            } else {
                x->nelems *= len;
            }
        }
        if (!err)
            err = skip_atts(nc);
        uint32_t t = 0;
        if (!err)
            err = rd_u32(nc, &t);
        if (!err && type_size((int)t) == 0)
            err = SWE_HOST_EFORMAT;
        x->type = (int)t;
        if (!err)
            err = rd_count(nc, wide, &x->vsize);
        if (!err)
            err = rd_count(nc, nc->version != 1, &x->begin);
        if (!err && (x->begin < 0 || x->begin > NC_MAX_BYTES || x->vsize < 0
                     || x->vsize > NC_MAX_BYTES))
            err = SWE_HOST_EFORMAT;
    }
    if (err) {
        nc_close_file(nc);
        return err == SWE_HOST_EREAD ? SWE_HOST_EFORMAT : err;
    }
    /* record size: the record variables' vsize summed; a single record
     * variable of a sub-4-byte type is not padded */
    int nrec = 0;
    for (long long v = 0; v < nc->nvars; v++)
        if (nc->vars[v].isrec) {
            nrec++;
            nc->recsize += nc->vars[v].vsize;   /* each < 2^62, at most 2^31 of them */
            if (nc->recsize > NC_MAX_BYTES) {
                nc_close_file(nc);
                return SWE_HOST_EFORMAT;
            }
        }
    if (nrec == 1)
        for (long long v = 0; v < nc->nvars; v++)
            if (nc->vars[v].isrec)
                nc->recsize = nc->vars[v].nelems * type_size(nc->vars[v].type);
    if (nc->numrecs < 0) {   /* streaming: records up to the end of the file */
        long long first = -1;
        for (long long v = 0; v < nc->nvars; v++)
            if (nc->vars[v].isrec && (first < 0 || nc->vars[v].begin < first))
                first = nc->vars[v].begin;
        nc->numrecs = 0;
        if (first >= 0 && nc->recsize > 0 && fseeko(nc->fp, 0, SEEK_END) == 0) {
            const long long end = (long long)ftello(nc->fp);
            nc->numrecs = end > first ? (end - first) / nc->recsize : 0;
        }
    }
    return 0;
}

static const ncvar *nc_find(const ncfile *nc, const char *name)
{
    for (long long v = 0; v < nc->nvars; v++)
        if (strcmp(nc->vars[v].name, name) == 0)
            return &nc->vars[v];
    return NULL;
}

/* Elements of x in the file; -1 when the count is past any file size. */
static long long var_count(const ncfile *nc, const ncvar *x)
{
    if (!x->isrec)
        return x->nelems;
    if (nc->numrecs > 0 && x->nelems > NC_MAX_BYTES / nc->numrecs)
        return -1;
    return x->nelems * nc->numrecs;
}

/* Element k of a big-endian buffer of type t as a double (integers of more
 * than 53 bits are rounded as a C cast rounds them). */
static double elem_f64(const unsigned char *p, int t)
{
    switch (t) {
    case NC_BYTE: return (double)(int8_t)p[0];
    case NC_UBYTE: return (double)p[0];
    case NC_SHORT: return (double)(int16_t)(uint16_t)be(p, 2);
    case NC_USHORT: return (double)(uint16_t)be(p, 2);
    case NC_INT: return (double)(int32_t)(uint32_t)be(p, 4);
    case NC_UINT: return (double)(uint32_t)be(p, 4);
    case NC_INT64: return (double)(int64_t)be(p, 8);
    case NC_UINT64: return (double)be(p, 8);
    case NC_FLOAT: {
        const uint32_t u = (uint32_t)be(p, 4);
        float f;
        memcpy(&f, &u, 4);
        return (double)f;
    }
    default: {
        const uint64_t u = be(p, 8);
        double d;
        memcpy(&d, &u, 8);
        return d;
    }
    }
}

/* Reads every element of x, converted by conv(p, type, out + k). */
typedef int (*conv_fn)(const unsigned char *p, int t, void *out, long long k);

static int conv_f64(const unsigned char *p, int t, void *out, long long k)
{
    ((double *)out)[k] = elem_f64(p, t);
    return 0;
}

/* nc_get_var_int: integers in int range as they are, floating values
 * truncated toward zero; anything outside int range is NC_ERANGE. */
static int conv_i32(const unsigned char *p, int t, void *out, long long k)
{
    if (t == NC_INT64 || t == NC_UINT64) {
        const uint64_t u = be(p, 8);
        const int in = t == NC_INT64 ? ((int64_t)u >= INT32_MIN && (int64_t)u <= INT32_MAX)
                                     : u <= (uint64_t)INT32_MAX;
        if (!in)
            return SWE_HOST_ERANGE;
        ((int *)out)[k] = (int)(int64_t)u;
        return 0;
    }
    const double d = elem_f64(p, t);
    if (!(d > -2147483649.0 && d < 2147483648.0))
        return SWE_HOST_ERANGE;
    ((int *)out)[k] = (int)d;
    return 0;
}

/* x holds exactly `want` numbers, all of them inside the file (checked
 * before anything is allocated for them). */
static int check_var(const ncfile *nc, const ncvar *x, long long want)
{
    if (x->type == NC_CHAR)
        return SWE_HOST_EFORMAT;   /* NC_ECHAR: text is not numeric */
    if (want < 0 || var_count(nc, x) != want)
        return SWE_HOST_EREAD;
    const long long es = type_size(x->type);
    const long long run = x->isrec ? x->nelems * es : want * es;
    if (x->isrec && nc->numrecs > 0
        && (nc->recsize <= 0 || nc->numrecs - 1 > nc->fsize / nc->recsize))
        return SWE_HOST_EREAD;   /* records past the end of the file */
    const long long last = x->isrec && nc->numrecs > 0 ? (nc->numrecs - 1) * nc->recsize : 0;
    if (x->begin + last > nc->fsize || run > nc->fsize - x->begin - last)
        return SWE_HOST_EREAD;   /* the data run past the end of the file */
    return 0;
}

static int read_var(ncfile *nc, const ncvar *x, long long want, conv_fn conv, void *out)
{
    const int bad = check_var(nc, x, want);
    if (bad)
        return bad;
    /* the records must lie inside the representable file range */
    if (x->isrec && nc->numrecs > 0
        && (nc->recsize > 0 && nc->numrecs - 1 > (NC_MAX_BYTES - x->begin) / nc->recsize))
        return SWE_HOST_EFORMAT;
    const int es = type_size(x->type);
    const long long per = x->isrec ? x->nelems : want;   /* contiguous elements per run */
    const long long runs = x->isrec ? nc->numrecs : 1;
    enum { CHUNK = 1 << 16 };
    unsigned char *buf = (unsigned char *)malloc((size_t)CHUNK * (size_t)es);
    if (!buf)
        return SWE_HOST_ENOMEM;
    int err = 0;
    long long k = 0;
    for (long long r = 0; r < runs && !err; r++) {
        if (fseeko(nc->fp, (off_t)(x->begin + r * nc->recsize), SEEK_SET) != 0) {
            err = SWE_HOST_EREAD;
            break;
        }
        for (long long done = 0; done < per && !err;) {
            const long long n = per - done < CHUNK ? per - done : CHUNK;
            if (rd(nc->fp, buf, (size_t)(n * es))) {
                err = SWE_HOST_EREAD;
                break;
            }
            for (long long q = 0; q < n && !err; q++)
                err = conv(buf + q * es, x->type, out, k + q);
            done += n;
            k += n;
        }
    }
    free(buf);
    return err;
}

static int get_f64(ncfile *nc, const char *name, long long count, double **out)
{
    *out = NULL;
    const ncvar *x = nc_find(nc, name);
    if (!x)
        return SWE_HOST_ENOVAR;
    const int bad = check_var(nc, x, count);
    if (bad)
        return bad;
    double *a = (double *)malloc(sizeof(double) * (size_t)(count ? count : 1));
    if (!a)
        return SWE_HOST_ENOMEM;
    const int err = read_var(nc, x, count, conv_f64, a);
    if (err) {
        free(a);
        return err;
    }
    *out = a;
    return 0;
}

static int get_scalar_f64(ncfile *nc, const char *name, double *out)
{
    const ncvar *x = nc_find(nc, name);
    return x ? read_var(nc, x, 1, conv_f64, out) : SWE_HOST_ENOVAR;
}

static int get_scalar_i32(ncfile *nc, const char *name, int *out)
{
    const ncvar *x = nc_find(nc, name);
    return x ? read_var(nc, x, 1, conv_i32, out) : SWE_HOST_ENOVAR;
}

static void round_f32(double *a, size_t count)
{
    for (size_t k = 0; k < count; k++)
        a[k] = (double)(float)a[k];
}

/* get_GSMD_nc, input.c:13-132. */
int swe_get_gsmd_nc(const char *path, swe_gsmd *g) { return swe_get_gsmd_nc_ex(path, 0, g); }

int swe_get_gsmd_nc_ex(const char *path, int fp32, swe_gsmd *g)
{
    memset(g, 0, sizeof(*g));
    ncfile nc;
    int err = nc_open_file(path, &nc);
    if (err)
        return err;
    double grav = 0.0;   /* read and checked as the reference does; not used */
    err = get_scalar_i32(&nc, "Nnodes", &g->Nnodes);
    if (!err) err = get_scalar_f64(&nc, "a", &g->a);
    if (!err) err = get_scalar_f64(&nc, "g", &grav);
    if (!err) err = get_scalar_f64(&nc, "gh0", &g->gh0);
    if (!err) err = get_scalar_f64(&nc, "timestep", &g->dt);
    if (!err && g->Nnodes <= 0)
        err = SWE_HOST_EREAD;
    const long long N = g->Nnodes;
    const char *sn[5] = {"x", "y", "z", "f", "ghm"};
    double **scal[5] = {&g->x, &g->y, &g->z, &g->f, &g->ghm};
    for (int k = 0; k < 5 && !err; k++)
        err = get_f64(&nc, sn[k], N, scal[k]);
    const char *vn[4] = {"p_u", "p_v", "p_w", "gradghm"};
    double **vec[4] = {&g->p_u, &g->p_v, &g->p_w, &g->gradghm};
    for (int k = 0; k < 4 && !err; k++)
        err = get_f64(&nc, vn[k], 3 * N, vec[k]);
    if (!err) err = get_scalar_i32(&nc, "Nnbr", &g->Nnbr);
    if (!err) err = get_scalar_f64(&nc, "gamma", &g->gamma);
    if (!err && g->Nnbr <= 0)
        err = SWE_HOST_EREAD;
    const long long NW = N * (long long)g->Nnbr;
    const char *dn[4] = {"DPx", "DPy", "DPz", "L"};
    double **dm[4] = {&g->Dx, &g->Dy, &g->Dz, &g->L};
    for (int k = 0; k < 4 && !err; k++)
        err = get_f64(&nc, dn[k], NW, dm[k]);
    if (!err) {
        const ncvar *x = nc_find(&nc, "idx");
        if (!x)
            err = SWE_HOST_ENOVAR;
        else if ((err = check_var(&nc, x, NW)) != 0)
            ;
        else if (!(g->idx = (int *)malloc(sizeof(int) * (size_t)NW)))
            err = SWE_HOST_ENOMEM;
        else
            err = read_var(&nc, x, NW, conv_i32, g->idx);
    }
    nc_close_file(&nc);
    if (err) {
        swe_gsmd_free(g);
        return err;
    }
    (void)grav;
    /* the FP32 model type: each value rounded to float as the binary
     * reader's FP32 path does (the reference's RUN_32BIT build reads these
     * variables with nc_get_var_double into float arrays, which is not
     * defined) */
    if (fp32) {
        for (int k = 0; k < 5; k++)
            round_f32(*scal[k], (size_t)N);
        for (int k = 0; k < 4; k++)
            round_f32(*vec[k], 3 * (size_t)N);
        for (int k = 0; k < 4; k++)
            round_f32(*dm[k], (size_t)NW);
    }
    /* unit sphere -> Earth radius, hyperviscosity (input.c:114-119) */
    for (long long k = 0; k < NW; k++) {
        g->Dx[k] = g->Dx[k] / g->a;
        g->Dy[k] = g->Dy[k] / g->a;
        g->Dz[k] = g->Dz[k] / g->a;
        g->L[k] = g->L[k] * g->gamma;
    }
    if (fp32) {
        for (int k = 0; k < 4; k++)
            round_f32(*dm[k], (size_t)NW);
        double *cst[4] = {&g->a, &g->gh0, &g->dt, &g->gamma};
        for (int k = 0; k < 4; k++)
            round_f32(cst[k], 1);
    }
    /* 1-based -> 0-based (input.c:127-129); ids outside the mesh rejected */
    for (long long k = 0; k < NW; k++) {
        g->idx[k] -= 1;
        if (g->idx[k] < 0 || g->idx[k] >= g->Nnodes) {
            swe_gsmd_free(g);
            return SWE_HOST_EREAD;
        }
    }
    g->padded_Nnodes = g->Nnodes;
    g->padded_Nnbr = g->Nnbr;
    return 0;
}

/* get_ICs_nc / get_FCs_nc (input.c:135-185): H or H_100, [Nnodes][4]. */
int swe_get_state_nc_ex(const char *path, int which, int fp32, double **H)
{
    *H = NULL;
    ncfile nc;
    int err = nc_open_file(path, &nc);
    if (err)
        return err;
    int N = 0;
    err = get_scalar_i32(&nc, "Nnodes", &N);
    if (!err && N <= 0)
        err = SWE_HOST_EREAD;
    if (!err)
        err = get_f64(&nc, which ? "H_100" : "H", 4LL * N, H);
    nc_close_file(&nc);
    if (!err && fp32)
        round_f32(*H, 4 * (size_t)N);
    return err;
}

/* nc2bin.c:10-84: the binary input file of the same case; the weights are
 * read (scaled by 1/a and gamma as get_GSMD_nc does), then unscaled again
 * (w * a, L / gamma) and the ids made 1-based before they are written. */
int swe_nc2bin(const char *nc_path, const char *bin_path)
{
    swe_gsmd g;
    int err = swe_get_gsmd_nc(nc_path, &g);
    if (err)
        return err;
    double *H0 = NULL, *H100 = NULL;
    err = swe_get_state_nc_ex(nc_path, 0, 0, &H0);
    if (!err)
        err = swe_get_state_nc_ex(nc_path, 1, 0, &H100);
    FILE *fp = NULL;
    if (!err && !(fp = fopen(bin_path, "wb")))
        err = SWE_HOST_EOPEN;
    if (!err) {
        const size_t N = (size_t)g.Nnodes, NW = N * (size_t)g.Nnbr;
        for (size_t k = 0; k < NW; k++) {
            g.Dx[k] = g.Dx[k] * g.a;
            g.Dy[k] = g.Dy[k] * g.a;
            g.Dz[k] = g.Dz[k] * g.a;
            g.L[k] = g.L[k] / g.gamma;
            g.idx[k] += 1;
        }
        const int32_t hdr[2] = {g.Nnodes, g.Nnbr};
        const double cst[4] = {g.a, g.gh0, g.dt, g.gamma};
        int ok = fwrite(hdr, sizeof(int32_t), 2, fp) == 2 && fwrite(cst, sizeof(double), 4, fp) == 4;
        const double *sc[5] = {g.x, g.y, g.z, g.f, g.ghm};
        for (int k = 0; k < 5 && ok; k++)
            ok = fwrite(sc[k], sizeof(double), N, fp) == N;
        const double *ve[4] = {g.p_u, g.p_v, g.p_w, g.gradghm};
        for (int k = 0; k < 4 && ok; k++)
            ok = fwrite(ve[k], sizeof(double), 3 * N, fp) == 3 * N;
        const double *dm[4] = {g.Dx, g.Dy, g.Dz, g.L};
        for (int k = 0; k < 4 && ok; k++)
            ok = fwrite(dm[k], sizeof(double), NW, fp) == NW;
        if (ok)
            ok = fwrite(g.idx, sizeof(int32_t), NW, fp) == NW;
        if (ok)
            ok = fwrite(H0, sizeof(double), 4 * N, fp) == 4 * N
                 && fwrite(H100, sizeof(double), 4 * N, fp) == 4 * N;
        if (fclose(fp) != 0 || !ok)
            err = SWE_HOST_EREAD;
    }
    free(H0);
    free(H100);
    swe_gsmd_free(&g);
    return err;
}
