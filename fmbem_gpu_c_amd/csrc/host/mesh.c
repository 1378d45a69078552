/* mesh.c -- triangle meshes: OBJ reader (findNum / readOBJ, octree.c:563-619),
 * icosphere and perturbed-ellipsoid generators, outward re-winding. */
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../fmb_internal.h"

static int mesh_alloc(int nv, int nf, fmb_mesh **out) {
  fmb_mesh *m = (fmb_mesh *)calloc(1, sizeof(fmb_mesh));
  if (!m) return fmb_set_error(FMB_ERR_NOMEM, "mesh allocation failed");
  m->nv = nv;
  m->nf = nf;
  m->V = (double *)malloc(sizeof(double) * 3 * (nv > 0 ? nv : 1));
  m->F = (int *)malloc(sizeof(int) * 3 * (nf > 0 ? nf : 1));
  if (!m->V || !m->F) {
    free(m->V);
    free(m->F);
    free(m);
    return fmb_set_error(FMB_ERR_NOMEM, "mesh allocation failed");
  }
  *out = m;
  return FMB_OK;
}

void fmb_mesh_destroy(fmb_mesh *m) {
  if (!m) return;
  free(m->V);
  free(m->F);
  free(m);
}

int fmb_mesh_sizes(const fmb_mesh *m, int *nv, int *nf) {
  if (!m) return fmb_set_error(FMB_ERR_ARG, "null mesh");
  if (nv) *nv = m->nv;
  if (nf) *nf = m->nf;
  return FMB_OK;
}

int fmb_mesh_export(const fmb_mesh *m, double *V, int *F) {
  if (!m) return fmb_set_error(FMB_ERR_ARG, "null mesh");
  if (V) memcpy(V, m->V, sizeof(double) * 3 * m->nv);
  if (F) memcpy(F, m->F, sizeof(int) * 3 * m->nf);
  return FMB_OK;
}

int fmb_mesh_from_arrays(const double *V, int nv, const int *F, int nf,
                         fmb_mesh **out) {
  if (!V || !F || nv < 3 || nf < 1 || !out)
    return fmb_set_error(FMB_ERR_ARG, "invalid mesh arrays (nv=%d nf=%d)", nv, nf);
  for (int i = 0; i < 3 * nf; i++)
    if (F[i] < 0 || F[i] >= nv)
      return fmb_set_error(FMB_ERR_ARG, "face %d references vertex %d of %d",
                           i / 3, F[i], nv);
  int rc = mesh_alloc(nv, nf, out);
  if (rc) return rc;
  memcpy((*out)->V, V, sizeof(double) * 3 * nv);
  memcpy((*out)->F, F, sizeof(int) * 3 * nf);
  return FMB_OK;
}

/* A 'v' record is "v" followed by blank; vn/vt/vp records are not vertices
 * (the reference counts any line starting with 'v', octree.c:576-578). */
static int is_rec(const char *line, char c) {
  return line[0] == c && (line[1] == ' ' || line[1] == '\t');
}

/* fgets that rejects a line longer than the buffer (a split line would be
 * read as two records); returns 1 for a line, 0 at end of file, -1 for an
 * overlong line */
static int read_line(char *buf, int size, FILE *fp) {
  if (!fgets(buf, size, fp)) return 0;
  size_t n = strlen(buf);
  if (n == (size_t)size - 1 && buf[n - 1] != '\n' && !feof(fp)) return -1;
  return 1;
}

#define OBJ_LINE 4096
#define OBJ_MAX_POLY 256 /* vertices of one polygon record */

int fmb_obj_count(const char *path, int *nv, int *nf) {
  if (!path || !nv || !nf) return fmb_set_error(FMB_ERR_ARG, "fmb_obj_count: null argument");
  FILE *fp = fopen(path, "r");
  if (!fp) return fmb_set_error(FMB_ERR_IO, "cannot open %s", path);
  char line[OBJ_LINE];
  long long v = 0, f = 0;
  int lno = 0, r;
  while ((r = read_line(line, sizeof(line), fp)) > 0) {
    lno++;
    if (is_rec(line, 'v')) v++;
    if (is_rec(line, 'f')) {
      /* polygons are fan-triangulated */
      int nt = 0;
      char *p = line + 1;
      while (*p) {
        while (*p && isspace((unsigned char)*p)) p++;
        if (!*p) break;
        nt++;
        while (*p && !isspace((unsigned char)*p)) p++;
      }
      if (nt > OBJ_MAX_POLY) {
        fclose(fp);
        return fmb_set_error(FMB_ERR_IO, "%s:%d: polygon of %d vertices (at most %d)", path,
                             lno, nt, OBJ_MAX_POLY);
      }
      if (nt >= 3) f += nt - 2;
    }
  }
  fclose(fp);
  if (r < 0) return fmb_set_error(FMB_ERR_IO, "%s:%d: line longer than %d bytes", path,
                                  lno + 1, OBJ_LINE - 1);
  if (v > 0x7fffffff / 3 || f > 0x7fffffff / 3)
    return fmb_set_error(FMB_ERR_IO, "%s: too many records", path);
  *nv = (int)v;
  *nf = (int)f;
  return FMB_OK;
}

/* fails the load: closes the file, frees the mesh, records the message */
#define OBJ_FAIL(...)                                  \
  do {                                                 \
    fclose(fp);                                        \
    fmb_mesh_destroy(m);                               \
    return fmb_set_error(FMB_ERR_IO, __VA_ARGS__);     \
  } while (0)

int fmb_mesh_load_obj(const char *path, fmb_mesh **out) {
  if (!out) return fmb_set_error(FMB_ERR_ARG, "fmb_mesh_load_obj: null output");
  int nv, nf, rc = fmb_obj_count(path, &nv, &nf);
  if (rc) return rc;
  if (nv < 3 || nf < 1)
    return fmb_set_error(FMB_ERR_IO, "%s: %d vertices, %d faces", path, nv, nf);
  fmb_mesh *m;
  if ((rc = mesh_alloc(nv, nf, &m))) return rc;
  FILE *fp = fopen(path, "r");
  if (!fp) {
    fmb_mesh_destroy(m);
    return fmb_set_error(FMB_ERR_IO, "cannot open %s", path);
  }
  char line[OBJ_LINE];
  int iv = 0, jf = 0, lno = 0, r;
  while ((r = read_line(line, sizeof(line), fp)) > 0) {
    lno++;
    if (is_rec(line, 'v')) {
      double x, y, z;
      if (iv >= nv) OBJ_FAIL("%s:%d: more vertices than counted (file changed)", path, lno);
      if (sscanf(line + 1, "%lf %lf %lf", &x, &y, &z) != 3)
        OBJ_FAIL("%s:%d: malformed vertex", path, lno);
      m->V[3 * iv] = x;
      m->V[3 * iv + 1] = y;
      m->V[3 * iv + 2] = z;
      iv++;
    } else if (is_rec(line, 'f')) {
      int ids[OBJ_MAX_POLY], n = 0;
      char *p = line + 1;
      while (*p) {
        while (*p && isspace((unsigned char)*p)) p++;
        if (!*p) break;
        if (n == OBJ_MAX_POLY)
          OBJ_FAIL("%s:%d: polygon of more than %d vertices", path, lno, OBJ_MAX_POLY);
        char *end;
        long v = strtol(p, &end, 10);
        /* 1-based ids (readOBJ, octree.c:606-609); negative = relative to
         * the vertices read so far; 0 is not an id */
        if (end == p || v == 0) OBJ_FAIL("%s:%d: malformed face", path, lno);
        long id = v > 0 ? v - 1 : iv + v;
        if (id < 0 || id >= nv)
          OBJ_FAIL("%s:%d: face references vertex %ld of %d", path, lno, v, nv);
        ids[n++] = (int)id;
        while (*p && !isspace((unsigned char)*p)) p++; /* skip /vt/vn */
      }
      if (n >= 3 && jf + (n - 2) > nf)
        OBJ_FAIL("%s:%d: more faces than counted (file changed)", path, lno);
      for (int t = 1; t + 1 < n; t++) {
        m->F[3 * jf] = ids[0];
        m->F[3 * jf + 1] = ids[t];
        m->F[3 * jf + 2] = ids[t + 1];
        jf++;
      }
    }
  }
  if (r < 0) OBJ_FAIL("%s:%d: line longer than %d bytes", path, lno + 1, OBJ_LINE - 1);
  if (iv != nv || jf != nf)
    OBJ_FAIL("%s: %d vertices and %d faces read, %d and %d counted (file changed)", path, iv,
             jf, nv, nf);
  fclose(fp);
  *out = m;
  return FMB_OK;
}
#undef OBJ_FAIL

/* signed volume with the reference normal (n0-n2)x(n1-n2) */
static double signed_volume(const fmb_mesh *m) {
  double vol = 0.0;
  for (int f = 0; f < m->nf; f++) {
    const double *a = m->V + 3 * m->F[3 * f], *b = m->V + 3 * m->F[3 * f + 1],
                 *c = m->V + 3 * m->F[3 * f + 2];
    double u[3], v[3], n[3];
    for (int q = 0; q < 3; q++) {
      u[q] = a[q] - c[q];
      v[q] = b[q] - c[q];
    }
    n[0] = u[1] * v[2] - u[2] * v[1];
    n[1] = u[2] * v[0] - u[0] * v[2];
    n[2] = u[0] * v[1] - u[1] * v[0];
    vol += (c[0] * n[0] + c[1] * n[1] + c[2] * n[2]) / 6.0;
  }
  return vol;
}

int fmb_mesh_orient_outward(fmb_mesh *m, int *nflipped) {
  if (!m) return fmb_set_error(FMB_ERR_ARG, "null mesh");
  int nfl = 0;
  if (signed_volume(m) < 0) {
    for (int f = 0; f < m->nf; f++) {
      int t = m->F[3 * f];
      m->F[3 * f] = m->F[3 * f + 1];
      m->F[3 * f + 1] = t;
    }
    nfl = m->nf;
  }
  if (nflipped) *nflipped = nfl;
  return FMB_OK;
}

/* ------------------------------------------------------------ icosphere */

typedef struct {
  double *V;
  int nv, cap;
} vbuf;

/* the callers size the buffer exactly (12 icosahedron vertices, then
 * 10 4^level + 2 for a subdivided icosphere) */
static int vpush(vbuf *b, const double *p) {
  if (b->nv == b->cap) return -1;
  memcpy(b->V + 3 * b->nv, p, sizeof(double) * 3);
  return b->nv++;
}

/* icosahedron with vertices at both poles, lower ring at azimuths
 * -36 - 72k degrees, upper ring at -72 - 72k degrees (the orientation of
 * the reference's Meshlab spheres). */
static void icosahedron(vbuf *vb, int *F) {
  double s5 = 1.0 / sqrt(5.0);
  double p[3] = {0, 0, -1};
  vpush(vb, p);
  for (int k = 0; k < 5; k++) {
    double a = (-36.0 - 72.0 * k) * M_PI / 180.0;
    double q[3] = {2 * s5 * cos(a), 2 * s5 * sin(a), -s5};
    vpush(vb, q);
  }
  for (int k = 0; k < 5; k++) {
    double a = (-72.0 - 72.0 * k) * M_PI / 180.0;
    double q[3] = {2 * s5 * cos(a), 2 * s5 * sin(a), s5};
    vpush(vb, q);
  }
  double t[3] = {0, 0, 1};
  vpush(vb, t);
  int n = 0;
  for (int k = 0; k < 5; k++) {
    int lo0 = 1 + k, lo1 = 1 + (k + 1) % 5, up0 = 6 + k, up1 = 6 + (k + 1) % 5;
    int f[4][3] = {{0, lo1, lo0}, {11, up0, up1}, {lo0, lo1, up0}, {up0, lo1, up1}};
    for (int q = 0; q < 4; q++)
      for (int c = 0; c < 3; c++) F[3 * n + 3 * q + c] = f[q][c];
    n += 4;
  }
}

/* edge midpoint cache: open addressing on (min,max) */
typedef struct {
  long long *key;
  int *val;
  size_t cap;
} emap;

static int emap_get(emap *m, int a, int b, vbuf *vb) {
  if (a > b) {
    int t = a;
    a = b;
    b = t;
  }
  long long k = ((long long)a << 32) | (unsigned)b;
  size_t h = (size_t)(k * 0x9E3779B97F4A7C15ULL) & (m->cap - 1);
  while (m->key[h] != -1) {
    if (m->key[h] == k) return m->val[h];
    h = (h + 1) & (m->cap - 1);
  }
  /* flat midpoint; all vertices are projected to the sphere once at the end,
   * which reproduces the vertex set of the reference's Meshlab spheres */
  double p[3];
  for (int q = 0; q < 3; q++) p[q] = 0.5 * (vb->V[3 * a + q] + vb->V[3 * b + q]);
  int id = vpush(vb, p);
  m->key[h] = k;
  m->val[h] = id;
  return id;
}

static void orient_faces_radially(const double *V, int *F, int nf) {
  for (int f = 0; f < nf; f++) {
    const double *a = V + 3 * F[3 * f], *b = V + 3 * F[3 * f + 1],
                 *c = V + 3 * F[3 * f + 2];
    double u[3], v[3], n[3], ct[3];
    for (int q = 0; q < 3; q++) {
      u[q] = a[q] - c[q];
      v[q] = b[q] - c[q];
      ct[q] = a[q] + b[q] + c[q];
    }
    n[0] = u[1] * v[2] - u[2] * v[1];
    n[1] = u[2] * v[0] - u[0] * v[2];
    n[2] = u[0] * v[1] - u[1] * v[0];
    if (n[0] * ct[0] + n[1] * ct[1] + n[2] * ct[2] < 0) {
      int t = F[3 * f];
      F[3 * f] = F[3 * f + 1];
      F[3 * f + 1] = t;
    }
  }
}

int fmb_mesh_icosphere(int level, double radius, fmb_mesh **out) {
  if (level < 0 || level > 10 || !(radius > 0) || !out)
    return fmb_set_error(FMB_ERR_ARG, "icosphere level %d radius %g", level, radius);
  /* 10 4^level + 2 vertices in all: the buffer never grows */
  vbuf vb = {0};
  vb.cap = 10 * (1 << (2 * level)) + 2;
  vb.V = (double *)malloc(sizeof(double) * 3 * (size_t)vb.cap);
  int nf = 20;
  int *F = (int *)malloc(sizeof(int) * 3 * 20);
  if (!vb.V || !F) {
    free(vb.V);
    free(F);
    return fmb_set_error(FMB_ERR_NOMEM, "icosphere level %d", level);
  }
  icosahedron(&vb, F);
  for (int l = 0; l < level; l++) {
    emap em;
    em.cap = 1;
    while (em.cap < (size_t)nf * 4) em.cap <<= 1;
    em.key = (long long *)malloc(sizeof(long long) * em.cap);
    em.val = (int *)malloc(sizeof(int) * em.cap);
    int *G = (int *)malloc(sizeof(int) * 3 * 4 * (size_t)nf);
    if (!em.key || !em.val || !G) {
      free(em.key);
      free(em.val);
      free(G);
      free(F);
      free(vb.V);
      return fmb_set_error(FMB_ERR_NOMEM, "icosphere level %d", level);
    }
    for (size_t i = 0; i < em.cap; i++) em.key[i] = -1;
    for (int f = 0; f < nf; f++) {
      int a = F[3 * f], b = F[3 * f + 1], c = F[3 * f + 2];
      int ab = emap_get(&em, a, b, &vb), bc = emap_get(&em, b, c, &vb),
          ca = emap_get(&em, c, a, &vb);
      int t[4][3] = {{a, ab, ca}, {b, bc, ab}, {c, ca, bc}, {ab, bc, ca}};
      for (int q = 0; q < 4; q++)
        for (int k = 0; k < 3; k++) G[12 * f + 3 * q + k] = t[q][k];
    }
    free(F);
    F = G;
    nf *= 4;
    free(em.key);
    free(em.val);
  }
  for (int i = 0; i < vb.nv; i++) {
    double *p = vb.V + 3 * i;
    double n = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    for (int q = 0; q < 3; q++) p[q] *= radius / n;
  }
  orient_faces_radially(vb.V, F, nf);
  fmb_mesh *m = (fmb_mesh *)calloc(1, sizeof(fmb_mesh));
  if (!m) {
    free(F);
    free(vb.V);
    return fmb_set_error(FMB_ERR_NOMEM, "icosphere level %d", level);
  }
  m->nv = vb.nv;
  m->nf = nf;
  m->V = vb.V;
  m->F = F;
  *out = m;
  return FMB_OK;
}

/* ------------------------------------------------------------ ellipsoid */

static uint64_t sm64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static double gauss01(uint64_t *s) {
  double u1 = ((sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  double u2 = ((sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* real spherical harmonic basis value (orthonormal), degree l, order m */
static double real_sh(int l, int m, double ct, double phi) {
  int am = m < 0 ? -m : m;
  /* normalised associated Legendre, no Condon-Shortley phase */
  double st = sqrt(fmax(0.0, 1.0 - ct * ct));
  double pmm = 1.0 / sqrt(4.0 * M_PI);
  for (int i = 1; i <= am; i++) pmm *= sqrt((2.0 * i + 1.0) / (2.0 * i)) * st;
  double pl = pmm;
  if (l > am) {
    double p1 = sqrt(2.0 * am + 3.0) * ct * pmm, p2 = pmm;
    pl = p1;
    for (int n = am + 2; n <= l; n++) {
      double a = sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)am * am));
      double b = sqrt(((double)(n - 1) * (n - 1) - (double)am * am) /
                      (4.0 * (n - 1) * (n - 1) - 1.0));
      double pn = a * (ct * p1 - b * p2);
      p2 = p1;
      p1 = pn;
      pl = pn;
    }
  }
  if (m == 0) return pl;
  return sqrt(2.0) * pl * (m > 0 ? cos(am * phi) : sin(am * phi));
}

int fmb_mesh_ellipsoid(int nu, double a, double b, double c, int lmin,
                       int lmax, double rms_frac, unsigned long long seed,
                       fmb_mesh **out) {
  if (nu < 1 || nu > 2000 || !(a > 0 && b > 0 && c > 0) || lmin < 0 ||
      lmax < lmin || lmax > 64 || !out)
    return fmb_set_error(FMB_ERR_ARG, "ellipsoid parameters");
  /* geodesic sphere: each icosahedron face split into nu^2 triangles */
  vbuf ib = {0};
  ib.cap = 12;
  ib.V = (double *)malloc(sizeof(double) * 3 * 12);
  if (!ib.V) return fmb_set_error(FMB_ERR_NOMEM, "ellipsoid mesh");
  int IF[60];
  icosahedron(&ib, IF);
  int nvmax = 10 * nu * nu + 2;
  double *V = (double *)malloc(sizeof(double) * 3 * nvmax);
  int nv = 0, nf = 20 * nu * nu, fi = 0;
  int *F = (int *)malloc(sizeof(int) * 3 * nf);
  /* vertex dedup through a hash on quantised positions */
  size_t cap = 1;
  while (cap < (size_t)nvmax * 4) cap <<= 1;
  long long *hk = (long long *)malloc(sizeof(long long) * cap * 3);
  int *hv = (int *)malloc(sizeof(int) * cap);
  int *grid = (int *)malloc(sizeof(int) * (nu + 1) * (nu + 1));
  if (!V || !F || !hk || !hv || !grid) {
    free(V);
    free(F);
    free(hk);
    free(hv);
    free(grid);
    free(ib.V);
    return fmb_set_error(FMB_ERR_NOMEM, "ellipsoid mesh, %d subdivisions", nu);
  }
  for (size_t i = 0; i < cap; i++) hv[i] = -1;
  for (int f = 0; f < 20; f++) {
    const double *A = ib.V + 3 * IF[3 * f], *B = ib.V + 3 * IF[3 * f + 1],
                 *C = ib.V + 3 * IF[3 * f + 2];
    for (int i = 0; i <= nu; i++)
      for (int j = 0; j <= nu - i; j++) {
        double p[3], n = 0;
        for (int q = 0; q < 3; q++) {
          p[q] = (A[q] * (nu - i - j) + B[q] * i + C[q] * j) / nu;
          n += p[q] * p[q];
        }
        n = sqrt(n);
        long long k3[3];
        for (int q = 0; q < 3; q++) {
          p[q] /= n;
          k3[q] = llround(p[q] * 1e9);
        }
        size_t h = (size_t)((k3[0] * 73856093LL) ^ (k3[1] * 19349663LL) ^
                            (k3[2] * 83492791LL)) &
                   (cap - 1);
        int id = -1;
        while (hv[h] != -1) {
          if (hk[3 * h] == k3[0] && hk[3 * h + 1] == k3[1] && hk[3 * h + 2] == k3[2]) {
            id = hv[h];
            break;
          }
          h = (h + 1) & (cap - 1);
        }
        if (id < 0) {
          id = nv++;
          memcpy(V + 3 * id, p, sizeof(double) * 3);
          hv[h] = id;
          hk[3 * h] = k3[0];
          hk[3 * h + 1] = k3[1];
          hk[3 * h + 2] = k3[2];
        }
        grid[i * (nu + 1) + j] = id;
      }
    for (int i = 0; i < nu; i++)
      for (int j = 0; j < nu - i; j++) {
        int v0 = grid[i * (nu + 1) + j], v1 = grid[(i + 1) * (nu + 1) + j],
            v2 = grid[i * (nu + 1) + j + 1];
        F[3 * fi] = v0;
        F[3 * fi + 1] = v1;
        F[3 * fi + 2] = v2;
        fi++;
        if (j + 1 < nu - i) {
          int v3 = grid[(i + 1) * (nu + 1) + j + 1];
          F[3 * fi] = v1;
          F[3 * fi + 1] = v3;
          F[3 * fi + 2] = v2;
          fi++;
        }
      }
  }
  free(grid);
  free(hk);
  free(hv);
  free(ib.V);
  orient_faces_radially(V, F, fi);
  /* radial perturbation field */
  uint64_t s = seed;
  int ncoef = (lmax + 1) * (lmax + 1);
  double *cf = (double *)calloc(ncoef, sizeof(double));
  for (int l = lmin; l <= lmax; l++)
    for (int m = -l; m <= l; m++) cf[l * l + l + m] = gauss01(&s);
  double *fv = (double *)malloc(sizeof(double) * nv);
  if (!cf || !fv) {
    free(cf);
    free(fv);
    free(V);
    free(F);
    return fmb_set_error(FMB_ERR_NOMEM, "ellipsoid perturbation");
  }
  double ss = 0;
  for (int i = 0; i < nv; i++) {
    double *p = V + 3 * i;
    double ct = p[2], phi = atan2(p[1], p[0]), f = 0;
    for (int l = lmin; l <= lmax; l++)
      for (int m = -l; m <= l; m++) f += cf[l * l + l + m] * real_sh(l, m, ct, phi);
    fv[i] = f;
    ss += f * f;
  }
  double rms = sqrt(ss / nv), scale = rms > 0 ? rms_frac / rms : 0.0;
  for (int i = 0; i < nv; i++) {
    double *p = V + 3 * i, g = 1.0 + scale * fv[i];
    p[0] *= a * g;
    p[1] *= b * g;
    p[2] *= c * g;
  }
  free(cf);
  free(fv);
  fmb_mesh *m = (fmb_mesh *)calloc(1, sizeof(fmb_mesh));
  if (!m) {
    free(V);
    free(F);
    return fmb_set_error(FMB_ERR_NOMEM, "ellipsoid mesh");
  }
  m->nv = nv;
  m->nf = fi;
  m->V = V;
  m->F = F;
  int nfl;
  fmb_mesh_orient_outward(m, &nfl);
  *out = m;
  return FMB_OK;
}
