/*
 * ddgi_oracle.c -- CPU oracle for the DDGI probe update (TEST INFRASTRUCTURE).
 *
 * Plain-C restatement of AneRend's probe-update path; see ddgi_oracle.h for
 * the file list, the pinning of each function and the usage restriction.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fno-fast-math -fopenmp).
 */
#include "ddgi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* float3 helpers: every operation is a single IEEE op, left to right.       */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } f3;

static inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 muls(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline f3 divs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* GLSL normalize(v) = v / length(v); contract: length = sqrt(dot(v, v)) */
static inline f3 normalize(f3 v) { float l = sqrtf(dot(v, v)); return divs(v, l); }
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
static inline void st3(float* p, f3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

/* Planted mistakes, for the pin tests only (tests/test_oracle_mutants.py builds
 * copies of this file with -DORA_MUTANT=<id> and checks that the closed-form
 * pins fail on each).  The default build, and every library loaded by the
 * other tests, has ORA_MUTANT == 0 and none of them:
 *   1  Schlick-GGX k = roughness^2 / 2          (pbr_light.glsl:26 has (r+1)^2/8)
 *   2  Fresnel exponent 4                        (pbr_light.glsl:90 has 5)
 *   3  wrap weight without the +0.2              (pbr_light.glsl:214)
 *   4  corners (1,1,0) and (1,0,0) read each other's tiles (pbr_light.glsl:326-335)
 *   5  tile start 10*i instead of 10*i + 1       (pbr_light.glsl:174-176)
 *   6  trilinear weights flipped: mix(alpha, 1 - alpha, offset) (pbr_light.glsl:194)
 *   7  indirect Fresnel with roughness, not roughness^2 (pbr_light.glsl:380)
 *   8  tile end start + 8 instead of start + 7   (pbr_light.glsl:178)
 *   9  GGX numerator without the +0.001          (pbr_light.glsl:17)
 *  10  probes blended in linear, not sqrt, space (pbr_light.glsl:235, 359)
 *  11  texel y/z swap left out of octEncode      (octahedron_helpers.glsl:10-13) */
#ifndef ORA_MUTANT
#define ORA_MUTANT 0
#endif
#define MUTANT(id) (ORA_MUTANT == (id))

static const float PI_F = 3.14159265359f;          /* pbr_light.glsl:8, rgen:14 */
static const double GOLDEN_RATIO_D = 1.6180339887; /* rgen:13 */
static const double PI_D = 3.14159265359;          /* rgen:14 */

/* ------------------------------------------------------------------------ */
/* Ray directions: rgen:35-46 (Fibonacci lattice) and rgen:87-88 (rotation). */
/* Contract: the lattice is evaluated in binary64 with libm, normalized in    */
/* binary64 and rounded once to binary32 (GLSL sin/cos/acos precision is     */
/* implementation-defined).  The rotation mat3(randomRot) * d is binary32.   */
/* ------------------------------------------------------------------------ */
void ora_fibonacci_dirs(int32_t n, float* out)
{
  for (int32_t i = 0; i < n; ++i) {
    double theta = 2.0 * PI_D * (double)i / GOLDEN_RATIO_D;
    double phi = acos(1.0 - 2.0 * ((double)i + 0.5) / (double)n);
    double x = cos(theta) * sin(phi);
    double y = sin(theta) * sin(phi);
    double z = cos(phi);
    double l = sqrt(x * x + y * y + z * z);
    out[3 * i + 0] = (float)(x / l);
    out[3 * i + 1] = (float)(y / l);
    out[3 * i + 2] = (float)(z / l);
  }
}

/* mat3(m) * v with m column-major: r[j] = m[0][j]*x + m[1][j]*y + m[2][j]*z */
static inline f3 rot3(const float* m, f3 v)
{
  return mk(m[0] * v.x + m[4] * v.y + m[8] * v.z,
            m[1] * v.x + m[5] * v.y + m[9] * v.z,
            m[2] * v.x + m[6] * v.y + m[10] * v.z);
}

void ora_rotate_dirs(int32_t n, const float* rot16, const float* in, float* out)
{
  for (int32_t i = 0; i < n; ++i) st3(out + 3 * i, rot3(rot16, ld3(in + 3 * i)));
}

/* ------------------------------------------------------------------------ */
/* Probe placement: rgen:60-84 (mirrored in pbr_light.glsl:294-308).         */
/* int() truncates toward zero (GLSL float->int conversion).                 */
/* ------------------------------------------------------------------------ */
static f3 cam_offset(const ora_config* c, const ora_frame* f)
{
  int cx = (int)(f->camera_pos[0] / c->probe_step[0]);
  int cz = (int)(f->camera_pos[2] / c->probe_step[2]);
  float ox = (float)cx * c->probe_step[0] - c->probe_step[0] * (float)c->probes_x / 2.0f;
  float oz = (float)cz * c->probe_step[2] - c->probe_step[2] * (float)c->probes_z / 2.0f;
  return mk(ox, 0.0f, oz);
}

void ora_probe_origin(const ora_config* c, const ora_frame* f, int32_t px, int32_t py,
                      int32_t pz, float* out)
{
  f3 p = mk((float)px, (float)py, (float)pz);
  f3 o;
  if (f->flags & ORA_FLAG_BAKE) {
    /* rgen:61-73 */
    int tsize = f->bake_tile[2];
    f3 scale = mk((float)tsize / (float)c->probes_x, 1.0f, (float)tsize / (float)c->probes_z);
    f3 toff = mk((float)f->bake_tile[0] * (float)tsize, 0.0f, (float)f->bake_tile[1] * (float)tsize);
    o = add(mul(p, scale), toff);
  } else {
    o = add(mul(p, ld3(c->probe_step)), cam_offset(c, f));
  }
  st3(out, o);
}

/* ------------------------------------------------------------------------ */
/* Octahedral mapping: octahedron_helpers.glsl:1-34 (y/z swapped).           */
/* ------------------------------------------------------------------------ */
static inline float sign_not_zero(float k) { return (k >= 0.0f) ? 1.0f : -1.0f; }

static void oct_encode(f3 v_in, float* rx, float* ry)
{
  f3 v = MUTANT(11) ? v_in : mk(v_in.x, v_in.z, v_in.y);
  float l1 = fabsf(v.x) + fabsf(v.y) + fabsf(v.z);
  float inv = 1.0f / l1;
  float x = v.x * inv, y = v.y * inv;
  if (v.z < 0.0f) {
    float nx = (1.0f - fabsf(y)) * sign_not_zero(x);
    float ny = (1.0f - fabsf(x)) * sign_not_zero(y);
    x = nx; y = ny;
  }
  *rx = x; *ry = y;
}

static f3 oct_decode(float ox, float oy)
{
  f3 v = mk(ox, oy, 1.0f - fabsf(ox) - fabsf(oy));
  if (v.z < 0.0f) {
    float nx = (1.0f - fabsf(v.y)) * sign_not_zero(v.x);
    float ny = (1.0f - fabsf(v.x)) * sign_not_zero(v.y);
    v.x = nx; v.y = ny;
  }
  return normalize(mk(v.x, v.z, v.y));
}

void ora_oct_encode(const float* v3, float* out2) { oct_encode(ld3(v3), out2, out2 + 1); }
void ora_oct_decode(const float* o2, float* out3) { st3(out3, oct_decode(o2[0], o2[1])); }

/* probe_conv.comp:37-46 with helpers.glsl:1-4: texel (tx, ty) of the 8x8 tile,
 * t = ty * 8 + tx. */
void ora_texel_dirs(float* out)
{
  for (int ty = 0; ty < 8; ++ty)
    for (int tx = 0; tx < 8; ++tx) {
      float u = (2.0f * (float)tx + 1.0f) / (2.0f * 8.0f);
      float v = (2.0f * (float)ty + 1.0f) / (2.0f * 8.0f);
      u = u * 2.0f - 1.0f;
      v = v * 2.0f - 1.0f;
      st3(out + 3 * (ty * 8 + tx), oct_decode(u, v));
    }
}

void ora_conv_weights(int32_t n, const float* dirs, float* w, float* wsum)
{
  float td[64 * 3];
  ora_texel_dirs(td);
  for (int t = 0; t < 64; ++t) {
    f3 o = ld3(td + 3 * t);
    float s = 0.0f;
    for (int r = 0; r < n; ++r) {
      float x = fmaxf(0.0f, dot(ld3(dirs + 3 * r), o));
      if (!(x > 1e-6f)) x = 0.0f;
      w[(int64_t)t * n + r] = x;
      s = s + x;
    }
    wsum[t] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* Scene and LBVH.                                                          */
/* ------------------------------------------------------------------------ */
struct ora_scene {
  /* inputs (copied) */
  float* vertices; int64_t nvert;
  uint32_t* indices; int64_t nidx;
  ora_mesh* meshes; int32_t nmesh;
  ora_material* mats; int32_t nmat;
  ora_instance* inst; int32_t ninst;
  float* inv_xform;   /* [ninst][12] world-to-object 3x4 row-major */
  /* flattened triangles, input order */
  int64_t ntri;
  int64_t* inst_tri_offset; /* [ninst + 1] */
  uint32_t* tri_inst;       /* [ntri] */
  float* world;             /* [ntri][9] */
  /* LBVH */
  uint64_t* keys;      /* sorted keys */
  uint32_t* ids;       /* sorted position -> triangle id */
  int32_t* left; int32_t* right; int32_t* parent; /* [ntri-1]; parent of leaves in leaf_parent */
  int32_t* leaf_parent;
  float* box;          /* [ntri-1][6] internal node boxes */
  float* leaf_box;     /* [ntri][6] in sorted order */
  int32_t max_depth;
  float* qbox;         /* [ntri-1][12] decoded quantized child boxes (null: exact boxes) */
};

static inline uint32_t mesh_of(const ora_instance* in) { return (in->custom_index_and_mask >> 10) & 0x3FFFu; }
static inline uint32_t mat_of(const ora_instance* in) { return in->custom_index_and_mask & 0x3FFu; }

/* Object-to-world with a 3x4 row-major matrix: r[j] = m[j][0]x + m[j][1]y + m[j][2]z + m[j][3].
 * Matches vec3(gl_ObjectToWorldEXT * vec4(p, 1)) (rchit:42). */
static inline f3 xform_point(const float* m, f3 p)
{
  return mk(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
            m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}

/* Affine inverse of a 3x4 row-major transform (host-side in the reference:
 * the driver derives gl_WorldToObjectEXT).  Contract: cofactor inverse in
 * binary64, rounded once to binary32. */
static void invert_3x4(const float* m, float* out)
{
  double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
  double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  double det = a * A + b * B + c * C;
  double inv[9];
  if (det == 0.0) det = 1e-30;
  inv[0] = A / det; inv[1] = -(b * i - c * h) / det; inv[2] = (b * f - c * e) / det;
  inv[3] = B / det; inv[4] = (a * i - c * g) / det;  inv[5] = -(a * f - c * d) / det;
  inv[6] = C / det; inv[7] = -(a * h - b * g) / det; inv[8] = (a * e - b * d) / det;
  double t0 = m[3], t1 = m[7], t2 = m[11];
  for (int r = 0; r < 3; ++r) {
    out[4 * r + 0] = (float)inv[3 * r + 0];
    out[4 * r + 1] = (float)inv[3 * r + 1];
    out[4 * r + 2] = (float)inv[3 * r + 2];
    out[4 * r + 3] = (float)(-(inv[3 * r + 0] * t0 + inv[3 * r + 1] * t1 + inv[3 * r + 2] * t2));
  }
}

static inline f3 vtx_pos(const ora_scene* s, uint32_t vi) { return ld3(s->vertices + (int64_t)vi * ORA_VERTEX_WORDS); }

static void flatten(ora_scene* s)
{
  for (int32_t k = 0; k < s->ninst; ++k) {
    const ora_instance* in = &s->inst[k];
    const ora_mesh* m = &s->meshes[mesh_of(in)];
    invert_3x4(in->transform, s->inv_xform + 12 * k);
    int64_t base = s->inst_tri_offset[k];
    int64_t nt = s->inst_tri_offset[k + 1] - base;
    for (int64_t p = 0; p < nt; ++p) {
      float* w = s->world + 9 * (base + p);
      for (int j = 0; j < 3; ++j) {
        uint32_t idx = s->indices[m->index_offset + 3 * p + j];
        f3 wp = xform_point(in->transform, vtx_pos(s, m->vertex_offset + idx));
        st3(w + 3 * j, wp);
      }
    }
  }
}

static inline void tri_box(const float* w, float* lo, float* hi)
{
  for (int a = 0; a < 3; ++a) {
    lo[a] = fminf(fminf(w[a], w[3 + a]), w[6 + a]);
    hi[a] = fmaxf(fmaxf(w[a], w[3 + a]), w[6 + a]);
  }
}

/* 21-bit spread for the 63-bit Morton code */
static inline uint64_t spread21(uint64_t v)
{
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

/* LBVH key contract: centroid c = (lo + hi) * 0.5 of the triangle box; the
 * centroid bounds [cmin, cmax] are mapped to a cube of side e = max extent
 * (aspect preserved); q = (c - cmin) / e * 2^21 clamped to [0, 2^21 - 1] and
 * truncated; key = x2 y1 z0 bit interleave (x most significant). */
static inline uint32_t quant21(float c, float cmin, float e)
{
  float q = (c - cmin) / e * 2097152.0f;
  q = fminf(fmaxf(q, 0.0f), 2097151.0f);
  return (uint32_t)q;
}

static int cmp_key(const void* a, const void* b)
{
  const uint64_t* x = (const uint64_t*)a;
  const uint64_t* y = (const uint64_t*)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  return 0;
}

/* delta(i, j): Karras 2012 section 4, keys augmented with the sorted index. */
static inline int delta(const uint64_t* k, int64_t n, int64_t i, int64_t j)
{
  if (j < 0 || j >= n) return -1;
  uint64_t a = k[i], b = k[j];
  if (a == b) return 64 + __builtin_clz((uint32_t)(i ^ j));
  return __builtin_clzll(a ^ b);
}

static void karras(ora_scene* s)
{
  int64_t n = s->ntri;
  const uint64_t* k = s->keys;
  for (int64_t i = 0; i < n - 1; ++i) {
    int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(k, n, i, i - d);
    int64_t lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
      if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = delta(k, n, i, j);
    int64_t sp = 0;
    int64_t div = 2;
    int64_t t = (l + div - 1) / div;
    for (;;) {
      if (delta(k, n, i, i + (sp + t) * d) > dnode) sp += t;
      if (t <= 1) break;
      div *= 2;
      t = (l + div - 1) / div;
    }
    int64_t gamma = i + sp * d + (d < 0 ? d : 0);
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    int32_t L = (lo == gamma) ? ~(int32_t)gamma : (int32_t)gamma;
    int32_t R = (hi == gamma + 1) ? ~(int32_t)(gamma + 1) : (int32_t)(gamma + 1);
    s->left[i] = L;
    s->right[i] = R;
    if (L < 0) s->leaf_parent[~L] = (int32_t)i; else s->parent[L] = (int32_t)i;
    if (R < 0) s->leaf_parent[~R] = (int32_t)i; else s->parent[R] = (int32_t)i;
  }
  if (n > 1) s->parent[0] = -1;
}

static void child_box(const ora_scene* s, int32_t c, float* lo, float* hi)
{
  const float* b = c < 0 ? s->leaf_box + 6 * (int64_t)(~c) : s->box + 6 * (int64_t)c;
  memcpy(lo, b, 12);
  memcpy(hi, b + 3, 12);
}

/* ------------------------------------------------------------------------ */
/* Quantized nodes (the library's 32-byte node layout, see                   */
/* ane_scene_set_node_format).  Node i keeps its box's lower corner o (binary32)  */
/* and per axis the smallest step 2^e (e >= -126) with 254 * 2^e >= the box   */
/* extent.  A child bound v is stored as q = floor((v - o) / 2^e) (lower) or   */
/* ceil((v - o) / 2^e) (upper), both exact in binary64, 0 <= q <= 254, and is  */
/* decoded in binary32 as o + q * 2^e (the product is exact, the sum rounds  */
/* to nearest).  Rounding is monotone, so the decoded box contains the exact  */
/* child box, and the traversal tests the decoded boxes.                      */
/* ------------------------------------------------------------------------ */
static int quant_exp(double extent)
{
  int e = -126;
  if (extent > 0.0) {
    int k;
    (void)frexp(extent, &k);  /* extent < 2^k */
    e = k - 9;                /* 254 * 2^(k-9) < 2^k */
    while (254.0 * ldexp(1.0, e) < extent) ++e;
    if (e < -126) e = -126;
  }
  return e;
}

static float quant_dec(float o, int q, int e) { return o + (float)q * ldexpf(1.0f, e); }

static void quantize_node(const float* l0, const float* h0, const float* l1, const float* h1, float* out12)
{
  for (int a = 0; a < 3; ++a) {
    const float o = l0[a] < l1[a] ? l0[a] : l1[a];
    const float hi = h0[a] > h1[a] ? h0[a] : h1[a];
    const int e = quant_exp((double)hi - (double)o);
    const double step = ldexp(1.0, e);
    out12[0 + a] = quant_dec(o, (int)floor(((double)l0[a] - (double)o) / step), e);
    out12[3 + a] = quant_dec(o, (int)ceil(((double)h0[a] - (double)o) / step), e);
    out12[6 + a] = quant_dec(o, (int)floor(((double)l1[a] - (double)o) / step), e);
    out12[9 + a] = quant_dec(o, (int)ceil(((double)h1[a] - (double)o) / step), e);
  }
}

static void quantize_nodes(ora_scene* s)
{
  if (!s->qbox || s->ntri < 2) return;
  for (int64_t i = 0; i < s->ntri - 1; ++i) {
    float l0[3], h0[3], l1[3], h1[3];
    child_box(s, s->left[i], l0, h0);
    child_box(s, s->right[i], l1, h1);
    quantize_node(l0, h0, l1, h1, s->qbox + 12 * i);
  }
}

void ora_scene_set_node32(ora_scene* s, int32_t on)
{
  free(s->qbox);
  s->qbox = NULL;
  if (on && s->ntri >= 2) {
    s->qbox = (float*)malloc(sizeof(float) * 12 * (size_t)(s->ntri - 1));
    quantize_nodes(s);
  }
}

void ora_scene_qboxes(const ora_scene* s, float* out)
{
  if (s->qbox) memcpy(out, s->qbox, sizeof(float) * 12 * (size_t)(s->ntri - 1));
}

/* Bottom-up boxes: parent box = componentwise min/max of child boxes (exact). */
static void fit_boxes(ora_scene* s)
{
  int64_t n = s->ntri;
  for (int64_t p = 0; p < n; ++p) {
    const float* w = s->world + 9 * (int64_t)s->ids[p];
    tri_box(w, s->leaf_box + 6 * p, s->leaf_box + 6 * p + 3);
  }
  if (n < 2) return;
  /* post-order iteration: children before parents */
  int32_t* st = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * n + 8));
  uint8_t* done = (uint8_t*)calloc((size_t)n, 1);
  int64_t sp = 0;
  st[sp++] = 0;
  while (sp) {
    int32_t i = st[sp - 1];
    int32_t L = s->left[i], R = s->right[i];
    if (!done[i]) {
      done[i] = 1;
      if (R >= 0) st[sp++] = R;
      if (L >= 0) st[sp++] = L;
      continue;
    }
    --sp;
    float l0[3], h0[3], l1[3], h1[3];
    child_box(s, L, l0, h0);
    child_box(s, R, l1, h1);
    float* b = s->box + 6 * (int64_t)i;
    for (int a = 0; a < 3; ++a) { b[a] = fminf(l0[a], l1[a]); b[3 + a] = fmaxf(h0[a], h1[a]); }
  }
  free(st);
  free(done);
}

static int32_t tree_depth(const ora_scene* s)
{
  int64_t n = s->ntri;
  if (n < 2) return 0;
  int32_t best = 0;
  for (int64_t p = 0; p < n; ++p) {
    int32_t d = 0;
    int32_t c = s->leaf_parent[p];
    while (c >= 0) { ++d; c = s->parent[c]; }
    if (d > best) best = d;
  }
  return best;
}

static void build_lbvh(ora_scene* s)
{
  int64_t n = s->ntri;
  if (n == 0) { s->max_depth = 0; return; }
  float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  float* cen = (float*)malloc(sizeof(float) * 3 * (size_t)n);
  for (int64_t t = 0; t < n; ++t) {
    float lo[3], hi[3];
    tri_box(s->world + 9 * t, lo, hi);
    for (int a = 0; a < 3; ++a) {
      float c = (lo[a] + hi[a]) * 0.5f;
      cen[3 * t + a] = c;
      cmin[a] = fminf(cmin[a], c);
      cmax[a] = fmaxf(cmax[a], c);
    }
  }
  float e = fmaxf(fmaxf(cmax[0] - cmin[0], cmax[1] - cmin[1]), cmax[2] - cmin[2]);
  if (!(e > 0.0f)) e = 1.0f;
  uint64_t* kv = (uint64_t*)malloc(sizeof(uint64_t) * 2 * (size_t)n);
  for (int64_t t = 0; t < n; ++t) {
    uint64_t x = quant21(cen[3 * t + 0], cmin[0], e);
    uint64_t y = quant21(cen[3 * t + 1], cmin[1], e);
    uint64_t z = quant21(cen[3 * t + 2], cmin[2], e);
    kv[2 * t] = (spread21(x) << 2) | (spread21(y) << 1) | spread21(z);
    kv[2 * t + 1] = (uint64_t)t;
  }
  qsort(kv, (size_t)n, 16, cmp_key); /* (key, id) lexicographic = stable sort by key */
  for (int64_t p = 0; p < n; ++p) { s->keys[p] = kv[2 * p]; s->ids[p] = (uint32_t)kv[2 * p + 1]; }
  free(kv);
  free(cen);
  karras(s);
  fit_boxes(s);
  s->max_depth = tree_depth(s);
}

ora_scene* ora_scene_create(const float* vertices, int64_t nvert, const uint32_t* indices,
                            int64_t nidx, const ora_mesh* meshes, int32_t nmesh,
                            const ora_material* mats, int32_t nmat, const ora_instance* inst,
                            int32_t ninst)
{
  ora_scene* s = (ora_scene*)calloc(1, sizeof(ora_scene));
  s->nvert = nvert; s->nidx = nidx; s->nmesh = nmesh; s->nmat = nmat; s->ninst = ninst;
  s->vertices = (float*)malloc(sizeof(float) * ORA_VERTEX_WORDS * (size_t)(nvert ? nvert : 1));
  memcpy(s->vertices, vertices, sizeof(float) * ORA_VERTEX_WORDS * (size_t)nvert);
  s->indices = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nidx ? nidx : 1));
  memcpy(s->indices, indices, sizeof(uint32_t) * (size_t)nidx);
  s->meshes = (ora_mesh*)malloc(sizeof(ora_mesh) * (size_t)(nmesh ? nmesh : 1));
  memcpy(s->meshes, meshes, sizeof(ora_mesh) * (size_t)nmesh);
  s->mats = (ora_material*)malloc(sizeof(ora_material) * (size_t)(nmat ? nmat : 1));
  memcpy(s->mats, mats, sizeof(ora_material) * (size_t)nmat);
  s->inst = (ora_instance*)malloc(sizeof(ora_instance) * (size_t)(ninst ? ninst : 1));
  memcpy(s->inst, inst, sizeof(ora_instance) * (size_t)ninst);
  s->inv_xform = (float*)malloc(sizeof(float) * 12 * (size_t)(ninst ? ninst : 1));
  s->inst_tri_offset = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ninst + 1));
  int64_t nt = 0;
  for (int32_t k = 0; k < ninst; ++k) {
    s->inst_tri_offset[k] = nt;
    nt += meshes[mesh_of(&inst[k])].index_count / 3;
  }
  s->inst_tri_offset[ninst] = nt;
  s->ntri = nt;
  size_t n1 = (size_t)(nt ? nt : 1);
  s->tri_inst = (uint32_t*)malloc(sizeof(uint32_t) * n1);
  for (int32_t k = 0; k < ninst; ++k)
    for (int64_t t = s->inst_tri_offset[k]; t < s->inst_tri_offset[k + 1]; ++t) s->tri_inst[t] = (uint32_t)k;
  s->world = (float*)malloc(sizeof(float) * 9 * n1);
  s->keys = (uint64_t*)malloc(sizeof(uint64_t) * n1);
  s->ids = (uint32_t*)malloc(sizeof(uint32_t) * n1);
  s->left = (int32_t*)malloc(sizeof(int32_t) * n1);
  s->right = (int32_t*)malloc(sizeof(int32_t) * n1);
  s->parent = (int32_t*)malloc(sizeof(int32_t) * n1);
  s->leaf_parent = (int32_t*)malloc(sizeof(int32_t) * n1);
  s->box = (float*)malloc(sizeof(float) * 6 * n1);
  s->leaf_box = (float*)malloc(sizeof(float) * 6 * n1);
  if (nt == 1) s->leaf_parent[0] = -1;
  flatten(s);
  build_lbvh(s);
  return s;
}

void ora_scene_destroy(ora_scene* s)
{
  if (!s) return;
  free(s->vertices); free(s->indices); free(s->meshes); free(s->mats); free(s->inst);
  free(s->inv_xform); free(s->inst_tri_offset); free(s->tri_inst); free(s->world);
  free(s->keys); free(s->ids); free(s->left); free(s->right); free(s->parent);
  free(s->leaf_parent); free(s->box); free(s->leaf_box); free(s->qbox);
  free(s);
}

/* Refit (UpdateBlasPass.cpp:131-138 MODE_UPDATE semantics): new transforms,
 * same topology, boxes recomputed bottom-up.  Instance meshes must not change. */
int ora_scene_refit(ora_scene* s, const ora_instance* inst, int32_t ninst)
{
  if (ninst != s->ninst) return -1;
  for (int32_t k = 0; k < ninst; ++k)
    if (mesh_of(&inst[k]) != mesh_of(&s->inst[k])) return -1;
  memcpy(s->inst, inst, sizeof(ora_instance) * (size_t)ninst);
  flatten(s);
  fit_boxes(s);
  quantize_nodes(s);
  return 0;
}

int64_t ora_scene_num_tris(const ora_scene* s) { return s->ntri; }
int32_t ora_scene_max_depth(const ora_scene* s) { return s->max_depth; }
void ora_scene_world_tris(const ora_scene* s, float* out) { memcpy(out, s->world, sizeof(float) * 9 * (size_t)s->ntri); }

void ora_scene_bvh(const ora_scene* s, uint64_t* keys, uint32_t* ids, int32_t* left,
                   int32_t* right, int32_t* parent, float* boxes)
{
  int64_t n = s->ntri;
  if (keys) memcpy(keys, s->keys, sizeof(uint64_t) * (size_t)n);
  if (ids) memcpy(ids, s->ids, sizeof(uint32_t) * (size_t)n);
  if (n < 2) return;
  if (left) memcpy(left, s->left, sizeof(int32_t) * (size_t)(n - 1));
  if (right) memcpy(right, s->right, sizeof(int32_t) * (size_t)(n - 1));
  if (parent) memcpy(parent, s->parent, sizeof(int32_t) * (size_t)(n - 1));
  if (boxes) memcpy(boxes, s->box, sizeof(float) * 6 * (size_t)(n - 1));
}

/* ------------------------------------------------------------------------ */
/* Ray / box and ray / triangle tests (contract shared with the HIP kernels) */
/* ------------------------------------------------------------------------ */
typedef struct {
  f3 o, d;
  f3 inv;           /* 1 / d with zero components replaced by +-1e-30 */
  int kx, ky, kz;
  float sx, sy, sz;
} ray_pre;

static inline float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

static void ray_setup(ray_pre* r, f3 o, f3 d)
{
  r->o = o; r->d = d;
  float dx = fabsf(d.x) < 1e-30f ? copysignf(1e-30f, d.x) : d.x;
  float dy = fabsf(d.y) < 1e-30f ? copysignf(1e-30f, d.y) : d.y;
  float dz = fabsf(d.z) < 1e-30f ? copysignf(1e-30f, d.z) : d.z;
  r->inv = mk(1.0f / dx, 1.0f / dy, 1.0f / dz);
  /* Woop et al. 2013, fig. 1: kz = dimension of max |d| */
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz + 1; if (kx == 3) kx = 0;
  int ky = kx + 1; if (ky == 3) ky = 0;
  if (comp(d, kz) < 0.0f) { int tmp = kx; kx = ky; ky = tmp; }
  r->kx = kx; r->ky = ky; r->kz = kz;
  r->sx = comp(d, kx) / comp(d, kz);
  r->sy = comp(d, ky) / comp(d, kz);
  r->sz = 1.0f / comp(d, kz);
}

/* Slab test, conservative per Ize 2013 (far distance scaled by 1 + 2*gamma(3),
 * rounded up to 1.0000004f).  tnear is clamped to tmin, tfar to tmax. */
#define ORA_ROBUST 1.0000004f
static inline int box_test(const ray_pre* r, const float* lo, const float* hi, float tmin,
                           float tmax, float* tnear)
{
  float tx0 = (lo[0] - r->o.x) * r->inv.x, tx1 = (hi[0] - r->o.x) * r->inv.x;
  float ty0 = (lo[1] - r->o.y) * r->inv.y, ty1 = (hi[1] - r->o.y) * r->inv.y;
  float tz0 = (lo[2] - r->o.z) * r->inv.z, tz1 = (hi[2] - r->o.z) * r->inv.z;
  float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
  float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  *tnear = tn;
  return tn <= tf * ORA_ROBUST;
}

/* Watertight test (Woop, Benthin, Wald 2013), no backface culling
 * (tlas_update.comp:86-87 sets the cull-disable flag).  Returns 1 and t, u, v
 * (u, v: barycentrics of vertices 1 and 2) if tmin <= t <= tmax, where
 * t = T * (1 / det), u = V * (1 / det), v = W * (1 / det). */
static inline int tri_test(const ray_pre* r, const float* w, float tmin, float tmax, float* t_out,
                           float* u_out, float* v_out)
{
  f3 A = sub(ld3(w), r->o), B = sub(ld3(w + 3), r->o), C = sub(ld3(w + 6), r->o);
  float Akz = comp(A, r->kz), Bkz = comp(B, r->kz), Ckz = comp(C, r->kz);
  float Ax = comp(A, r->kx) - r->sx * Akz, Ay = comp(A, r->ky) - r->sy * Akz;
  float Bx = comp(B, r->kx) - r->sx * Bkz, By = comp(B, r->ky) - r->sy * Bkz;
  float Cx = comp(C, r->kx) - r->sx * Ckz, Cy = comp(C, r->ky) - r->sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
    U = (float)(CxBy - CyBx);
    double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
    V = (float)(AxCy - AyCx);
    double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
    W = (float)(BxAy - ByAx);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
  float det = U + V + W;
  if (det == 0.0f) return 0;
  float Az = r->sz * Akz, Bz = r->sz * Bkz, Cz = r->sz * Ckz;
  float T = U * Az + V * Bz + W * Cz;
  /* contract: one correctly rounded reciprocal, then products */
  float rcp = 1.0f / det;
  float t = T * rcp;
  if (!(t >= tmin && t <= tmax)) return 0;
  *t_out = t;
  *u_out = V * rcp;
  *v_out = W * rcp;
  return 1;
}

/* Closest hit = lexicographic minimum of (t, triangle id) over all triangles
 * with tmin <= t <= tmax: independent of BVH shape and traversal order. */
static inline int better(float t, int32_t id, float bt, int32_t bid)
{
  return t < bt || (t == bt && id < bid);
}

#define ORA_STACK 128
static void trace_one(const ora_scene* s, f3 o, f3 d, float tmin, float tmax, int any,
                      ora_hit* out, uint64_t* nodes, uint64_t* tris)
{
  ray_pre r;
  ray_setup(&r, o, d);
  ora_hit best = {-1, tmax, 0.0f, 0.0f};
  uint64_t cn = 0, ct = 0;
  int64_t n = s->ntri;
  if (n == 0) goto done;
  /* a zero or non-finite direction is no ray */
  if ((d.x == 0.0f && d.y == 0.0f && d.z == 0.0f) || !isfinite(d.x) || !isfinite(d.y) || !isfinite(d.z))
    goto done;
  int32_t stack[ORA_STACK];
  int sp = 0;
  int32_t node = (n == 1) ? ~0 : 0;
  for (;;) {
    if (node >= 0) {
      ++cn;
      int32_t L = s->left[node], R = s->right[node];
      float l0[3], h0[3], l1[3], h1[3], tn0, tn1;
      if (s->qbox) {
        const float* q = s->qbox + 12 * (int64_t)node;
        memcpy(l0, q, 12); memcpy(h0, q + 3, 12); memcpy(l1, q + 6, 12); memcpy(h1, q + 9, 12);
      } else {
        child_box(s, L, l0, h0);
        child_box(s, R, l1, h1);
      }
      int b0 = box_test(&r, l0, h0, tmin, best.t, &tn0);
      int b1 = box_test(&r, l1, h1, tmin, best.t, &tn1);
      if (b0 || b1) {
        if (b0 && b1) {
          int32_t nearc = (tn1 < tn0) ? R : L;
          int32_t farc = (tn1 < tn0) ? L : R;
          stack[sp++] = farc;
          node = nearc;
        } else {
          node = b0 ? L : R;
        }
        continue;
      }
    } else {
      ++ct;
      int32_t id = (int32_t)s->ids[~node];
      float t, u, v;
      if (tri_test(&r, s->world + 9 * (int64_t)id, tmin, best.t, &t, &u, &v)) {
        if (any) { best.tri = id; best.t = t; best.u = u; best.v = v; break; }
        if (better(t, id, best.t, best.tri)) { best.tri = id; best.t = t; best.u = u; best.v = v; }
      }
    }
    if (sp == 0) break;
    node = stack[--sp];
  }
done:
  if (best.tri < 0) best.t = 0.0f;
  *out = best;
  if (nodes) *nodes += cn;
  if (tris) *tris += ct;
}

void ora_trace(const ora_scene* s, const float* orig, const float* dir, int64_t nrays,
               float tmin, float tmax, int32_t any_hit, ora_hit* out, uint64_t* counters)
{
  uint64_t cn = 0, ct = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : cn, ct)
  for (int64_t i = 0; i < nrays; ++i)
    trace_one(s, ld3(orig + 3 * i), ld3(dir + 3 * i), tmin, tmax, any_hit, out + i, &cn, &ct);
  if (counters) { counters[0] += cn; counters[1] += ct; }
}

/* shadow.rgen:29-82: visibility of each depth-buffer pixel toward the sun */
void ora_shadow_mask(const ora_scene* s, const float* depth, int32_t w, int32_t h, const float* ivp,
                     const float* light, float* out, uint64_t* counters)
{
  uint64_t cn = 0, ct = 0;
  f3 dir = mk(-light[0], -light[1], -light[2]);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : cn, ct)
  for (int64_t i = 0; i < (int64_t)w * h; ++i) {
    int32_t y = (int32_t)(i / w), x = (int32_t)(i - (int64_t)y * w);
    float d = depth[i];
    if (d == 1.0f) { out[i] = 1.0f; continue; }
    float u = (2.0f * (float)x + 1.0f) / (2.0f * (float)w);   /* pixelToUv, helpers.glsl:1-4 */
    float v = (2.0f * (float)y + 1.0f) / (2.0f * (float)h);
    float cx = u * 2.0f - 1.0f, cy = v * 2.0f - 1.0f;         /* calcWorldPos, shadow.rgen:20-27 */
    float px = ivp[0] * cx + ivp[4] * cy + ivp[8] * d + ivp[12];
    float py = ivp[1] * cx + ivp[5] * cy + ivp[9] * d + ivp[13];
    float pz = ivp[2] * cx + ivp[6] * cy + ivp[10] * d + ivp[14];
    float pw = ivp[3] * cx + ivp[7] * cy + ivp[11] * d + ivp[15];
    ora_hit hit;
    trace_one(s, mk(px / pw, py / pw, pz / pw), dir, 0.1f, 50.0f, 1, &hit, &cn, &ct);
    out[i] = hit.tri >= 0 ? 0.0f : 1.0f;
  }
  if (counters) { counters[0] += cn; counters[1] += ct; }
}

/* ------------------------------------------------------------------------ */
/* Point-light shadow moments: light_shadow.rgen:27-86, light_shadow.rchit:   */
/* 29-32, light_shadow.rmiss:12-15 (LightShadowRayTracingPass.cpp:17-100),    */
/* light_shadow_sum.comp:16-71 (LightShadowSumPass.cpp:16-87), constants of   */
/* light_shadow_helpers.glsl:1-4.                                             */
/* ------------------------------------------------------------------------ */

/* pow(x, 50.0) of light_shadow_sum.comp:58.  Contract (GLSL pow precision is
 * implementation-defined): the binary32 product chain x^2, x^4, x^8, x^16,
 * x^32, then (x^32 * x^16) * x^2. */
static float pow50(float x)
{
  float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16;
  return (x32 * x16) * x2;
}

/* The direction image is RGBA8_UNORM (LightShadowRayTracingPass.cpp:53): the
 * store of (dir + 1) * 0.5 (light_shadow.rgen:83-85) and the load
 * (light_shadow_sum.comp:56-57).  Contract: clamp to [0, 1], scale by 255,
 * round to nearest even; the load is c / 255. */
static float unorm8_round_trip(float v)
{
  float c = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  return rintf(c * 255.0f) / 255.0f;
}

void ora_light_shadow_weights(float* w)
{
  float dirs[3 * ORA_LIGHT_RAYS], td[64 * 3];
  ora_fibonacci_dirs(ORA_LIGHT_RAYS, dirs);  /* worldDirFromFibonacciIndex, rgen:27-38 */
  ora_texel_dirs(td);                        /* octUv / octDecode, sum.comp:45-47 */
  for (int r = 0; r < ORA_LIGHT_RAYS; ++r) {
    f3 q;
    q.x = unorm8_round_trip((dirs[3 * r + 0] + 1.0f) * 0.5f) * 2.0f - 1.0f;
    q.y = unorm8_round_trip((dirs[3 * r + 1] + 1.0f) * 0.5f) * 2.0f - 1.0f;
    q.z = unorm8_round_trip((dirs[3 * r + 2] + 1.0f) * 0.5f) * 2.0f - 1.0f;
    for (int t = 0; t < 64; ++t) {
      float c = dot(q, ld3(td + 3 * t));
      float wt = pow50(c > 0.0f ? c : 0.0f);  /* pow(max(0.0, dot(dir, octDir)), 50.0) */
      w[t * ORA_LIGHT_RAYS + r] = wt > 0.001f ? wt : 0.0f;  /* sum.comp:59 */
    }
  }
}

void ora_light_shadows(const ora_scene* s, const float* lights, int32_t n, float* dist_img,
                       float* moments_img, uint64_t* counters)
{
  float dirs[3 * ORA_LIGHT_RAYS];
  ora_fibonacci_dirs(ORA_LIGHT_RAYS, dirs);
  float* dist = (float*)malloc(sizeof(float) * (size_t)n * ORA_LIGHT_RAYS);
  float* w = (float*)malloc(sizeof(float) * 64 * ORA_LIGHT_RAYS);
  ora_light_shadow_weights(w);
  uint64_t cn = 0, ct = 0;
  const int64_t total = (int64_t)n * ORA_LIGHT_RAYS;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : cn, ct)
  for (int64_t i = 0; i < total; ++i) {
    const int32_t l = (int32_t)(i / ORA_LIGHT_RAYS), r = (int32_t)(i % ORA_LIGHT_RAYS);
    const float* L = lights + 8 * (int64_t)l;  /* {worldPos (w = range), color (w = enabled)} */
    float d;
    if (L[7] < 0.5f) {
      d = 0.0f;  /* disabled light: rgen:54-58 */
    } else {
      ora_hit h;
      trace_one(s, mk(L[0], L[1], L[2]), ld3(dirs + 3 * r), 0.1f, L[3], 0, &h, &cn, &ct);
      d = h.tri >= 0 ? h.t : L[3];  /* rchit: gl_HitTEXT; rmiss: maxDist = range */
    }
    dist[i] = d;
    /* rayIndex = y * 32 + x at pixel (32 * l + x, y) of the distance image (rgen:43-49) */
    if (dist_img)
      dist_img[(int64_t)(r / ORA_LIGHT_SIDE) * ORA_LIGHT_SIDE * n + ORA_LIGHT_SIDE * l + r % ORA_LIGHT_SIDE] = d;
  }
  const int64_t row = 10 * (int64_t)n;  /* (LIGHT_SHADOW_OCT_SIZE + 2) texels per light */
  for (int32_t l = 0; l < n; ++l)
    for (int t = 0; t < 64; ++t) {
      const int tx = t & 7, ty = t >> 3;
      float acc_r = 0.0f, acc_g = 0.0f, acc_w = 0.0f;
      if (!(lights[8 * (int64_t)l + 7] < 0.5f)) {  /* sum.comp:35: disabled when color.w < 0.5 */
        for (int x = 0; x < ORA_LIGHT_SIDE; ++x)      /* sum.comp:50-63: x outer, y inner */
          for (int y = 0; y < ORA_LIGHT_SIDE; ++y) {
            const int idx = y * ORA_LIGHT_SIDE + x;
            const float wt = w[t * ORA_LIGHT_RAYS + idx];
            if (wt > 0.001f) {
              const float d = dist[(int64_t)l * ORA_LIGHT_RAYS + idx];
              acc_r += wt * d;
              acc_g += wt * d * d;
              acc_w += wt;
            }
          }
        if (acc_w > 1e-6f) {
          acc_r = acc_r / acc_w;
          acc_g = acc_g / acc_w;
        }
      }
      /* texel (10 * l + 1 + tx, ty) of the RG32F moments image (sum.comp:29-31, 69) */
      float* o = moments_img + 2 * ((int64_t)ty * row + 10 * (int64_t)l + 1 + tx);
      o[0] = acc_r;
      o[1] = acc_g;
    }
  free(w);
  free(dist);
  if (counters) { counters[0] += cn; counters[1] += ct; }
}

void ora_trace_bruteforce(const ora_scene* s, const float* orig, const float* dir,
                          int64_t nrays, float tmin, float tmax, int32_t any_hit, ora_hit* out)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nrays; ++i) {
    ray_pre r;
    ray_setup(&r, ld3(orig + 3 * i), ld3(dir + 3 * i));
    ora_hit best = {-1, tmax, 0.0f, 0.0f};
    for (int64_t t = 0; t < s->ntri; ++t) {
      float th, u, v;
      if (tri_test(&r, s->world + 9 * t, tmin, best.t, &th, &u, &v)) {
        if (any_hit) { best.tri = (int32_t)t; best.t = th; best.u = u; best.v = v; break; }
        if (better(th, (int32_t)t, best.t, best.tri)) { best.tri = (int32_t)t; best.t = th; best.u = u; best.v = v; }
      }
    }
    if (best.tri < 0) best.t = 0.0f;
    out[i] = best;
  }
}

/* ------------------------------------------------------------------------ */
/* Shading: rchit:24-110, rmiss:17-21, pbr_light.glsl, bindless.glsl.        */
/* ------------------------------------------------------------------------ */
/* bindless.glsl:240-247; pow is libm powf (contract) */
static inline float to_linear1(float c)
{
  return (c < 0.04045f) ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

void ora_sky_radiance(const ora_frame* f, float* out)
{
  /* rmiss:20: toLinear(skyIntensity * skyColor) */
  for (int a = 0; a < 3; ++a) out[a] = to_linear1(f->sky_intensity * f->sky_color[a]);
}

/* pbr_light.glsl:10-22 */
static inline float distribution_ggx(f3 N, f3 H, float rough)
{
  float a = rough * rough;
  float a2 = a * a;
  float ndoth = fmaxf(dot(N, H), 0.0f);
  float ndoth2 = ndoth * ndoth;
  float num = MUTANT(9) ? a2 : a2 + 0.001f;
  float den = ndoth2 * (a2 - 1.0f) + 1.0f;
  den = PI_F * den * den;
  return num / den;
}

/* pbr_light.glsl:24-43 */
static inline float schlick_ggx(float ndotv, float rough)
{
  float r = rough + 1.0f;
  float k = MUTANT(1) ? rough * rough / 2.0f : (r * r) / 8.0f;
  return ndotv / (ndotv * (1.0f - k) + k);
}

/* pbr_light.glsl:88-91: pow(clamp(1 - c, 0, 1), 5) via libm powf (contract) */
static inline float pow5_term(float c)
{
  return powf(clampf(1.0f - c, 0.0f, 1.0f), MUTANT(2) ? 4.0f : 5.0f);
}

/* pbr_light.glsl:98-166 with directional == 1, useSpecular == true */
static f3 calc_light(const ora_frame* f, f3 N, f3 albedo, float metallic, float rough, f3 pos,
                     f3 view)
{
  f3 light_color = muls(mk(1.0f, 1.0f, 1.0f), f->sun_intensity);
  f3 V = normalize(sub(view, pos));
  f3 F0 = add(muls(mk(0.04f, 0.04f, 0.04f), 1.0f - metallic), muls(albedo, metallic));
  f3 L = normalize(mk(-f->light_dir[0], -f->light_dir[1], -f->light_dir[2]));
  f3 H = normalize(add(V, L));
  f3 radiance = muls(light_color, 1.0f);
  float ndf = distribution_ggx(N, H, rough);
  float ndotv_ = fmaxf(dot(N, V), 0.0f);
  float ndotl_ = fmaxf(dot(N, L), 0.0f);
  float G = schlick_ggx(ndotl_, rough) * schlick_ggx(ndotv_, rough);
  float p5 = pow5_term(fmaxf(dot(H, V), 0.0f));
  f3 F = add(F0, muls(sub(mk(1.0f, 1.0f, 1.0f), F0), p5));
  f3 kD = muls(sub(mk(1.0f, 1.0f, 1.0f), F), 1.0f - metallic);
  float ndotl = fmaxf(dot(N, L), 0.0f);
  float ndotv = fmaxf(dot(N, V), 0.0f);
  f3 numer = muls(F, ndf * G);
  float denom = 4.0f * ndotv * ndotl + 0.0001f;
  f3 spec = divs(numer, denom);
  spec = mk(clampf(spec.x, 0.0f, 1.0f), clampf(spec.y, 0.0f, 1.0f), clampf(spec.z, 0.0f, 1.0f));
  f3 lo = add(divs(mul(kD, albedo), PI_F), spec);
  return muls(mul(lo, radiance), ndotl);
}

/* Atlas geometry: IrradianceProbeConvolvePass.cpp:89-90 */
int64_t ora_atlas_width(const ora_config* c) { return 10 * (int64_t)c->probes_x; }
int64_t ora_atlas_height(const ora_config* c) { return 10 * (int64_t)c->probes_z * c->probes_y; }

void ora_atlas_clear(const ora_config* c, float* atlas)
{
  /* IrradianceProbeConvolvePass.cpp:20-24 clear to (0, 0, 0, 1) */
  int64_t n = ora_atlas_width(c) * ora_atlas_height(c);
  for (int64_t i = 0; i < n; ++i) {
    atlas[4 * i + 0] = 0.0f; atlas[4 * i + 1] = 0.0f; atlas[4 * i + 2] = 0.0f; atlas[4 * i + 3] = 1.0f;
  }
}

static inline int imod(int a, int m) { int r = a % m; return r < 0 ? r + m : r; }

/* texture(probeTex, uv) with LINEAR filter and REPEAT wrap (PipelineUtil.h:98-112).
 * Contract: binary32 weights (hardware uses fixed-point fractions). */
static f3 bilinear_repeat(const float* atlas, int W, int H, float u, float v)
{
  float x = u * (float)W - 0.5f;
  float y = v * (float)H - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float ax = x - fx, ay = y - fy;
  int x0 = imod((int)fx, W), x1 = imod((int)fx + 1, W);
  int y0 = imod((int)fy, H), y1 = imod((int)fy + 1, H);
  const float* t00 = atlas + 4 * ((int64_t)y0 * W + x0);
  const float* t10 = atlas + 4 * ((int64_t)y0 * W + x1);
  const float* t01 = atlas + 4 * ((int64_t)y1 * W + x0);
  const float* t11 = atlas + 4 * ((int64_t)y1 * W + x1);
  f3 r;
  float* rp = &r.x;
  for (int c = 0; c < 3; ++c) {
    float top = t00[c] * (1.0f - ax) + t10[c] * ax;
    float bot = t01[c] * (1.0f - ax) + t11[c] * ax;
    rp[c] = top * (1.0f - ay) + bot * ay;
  }
  return r;
}

/* pbr_light.glsl:168-188 */
static f3 sample_single_probe(const ora_config* c, const float* atlas, int ix, int iy, int iz, f3 n)
{
  int W = (int)ora_atlas_width(c), H = (int)ora_atlas_height(c);
  int sx = 10 * ix + (MUTANT(5) ? 0 : 1);
  int sy = 10 * c->probes_z * iy + 10 * iz + (MUTANT(5) ? 0 : 1);
  int ex = sx + (MUTANT(8) ? 8 : 7), ey = sy + (MUTANT(8) ? 8 : 7);
  float su = (2.0f * (float)sx + 1.0f) / (2.0f * (float)W);
  float sv = (2.0f * (float)sy + 1.0f) / (2.0f * (float)H);
  float eu = (2.0f * (float)ex + 1.0f) / (2.0f * (float)W);
  float ev = (2.0f * (float)ey + 1.0f) / (2.0f * (float)H);
  float ox, oy;
  oct_encode(normalize(n), &ox, &oy);
  ox = (ox + 1.0f) * 0.5f;
  oy = (oy + 1.0f) * 0.5f;
  float u = su + (eu - su) * ox;
  float v = sv + (ev - sv) * oy;
  return bilinear_repeat(atlas, W, H, u, v);
}

/* ------------------------------------------------------------------------ */
/* Baked tile atlases: Scene::setDDGIAtlas (Scene.cpp:409-438) registers a    */
/* baked RGBA16F atlas per world tile; prefillGPUTileInfoBuffer               */
/* (VulkanRenderer.cpp:3698-3738) writes the camera-relative window of        */
/* GPUTileInfo (GpuBuffers.h:74-76) the shaders read through                  */
/* worldPosToTileInfo (bindless.glsl:291-308).  Contract for Q14: the window  */
/* radius is ORA_TILE_RADIUS = 2 (MAX_PAGE_TILE_RADIUS, VulkanRenderer.h:267, */
/* the buffer the reference fills) on both sides, with floor() for the        */
/* camera tile as in the shader; beyond it the lookup returns -1.  Test-only  */
/* global state: the tiles hold their atlas widened to binary32.              */
/* ------------------------------------------------------------------------ */
#define ORA_TILE_RADIUS 2
static struct {
  int32_t n, size;
  int32_t* xz;
  float* atlas; /* [n][H][W][4] */
  int64_t texels;
} g_tiles = {0, 32, NULL, NULL, 0};

void ora_tiles_set(const ora_config* c, const int32_t* tile_xz, const float* atlases, int32_t n,
                   int32_t tile_size)
{
  free(g_tiles.xz);
  free(g_tiles.atlas);
  g_tiles.xz = NULL;
  g_tiles.atlas = NULL;
  g_tiles.n = 0;
  g_tiles.size = tile_size;
  if (n <= 0) return;
  int64_t texels = ora_atlas_width(c) * ora_atlas_height(c);
  g_tiles.xz = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)n);
  g_tiles.atlas = (float*)malloc(sizeof(float) * 4 * (size_t)(texels * n));
  memcpy(g_tiles.xz, tile_xz, sizeof(int32_t) * 2 * (size_t)n);
  memcpy(g_tiles.atlas, atlases, sizeof(float) * 4 * (size_t)(texels * n));
  g_tiles.texels = texels;
  g_tiles.n = n;
}

/* worldPosToTileInfo + tileInfoBuffer.tiles[i].ddgiAtlasTex: the baked atlas of
 * the tile holding pos, or NULL; *tx, *tz receive the tile index */
static const float* tile_atlas(const ora_frame* f, f3 pos, int* tx, int* tz)
{
  const float ts = (float)g_tiles.size;
  *tx = (int)floorf(pos.x / ts);
  *tz = (int)floorf(pos.z / ts);
  if (g_tiles.n == 0) return NULL;
  int cx = (int)floorf(f->camera_pos[0] / ts), cz = (int)floorf(f->camera_pos[2] / ts);
  int rx = *tx - cx, rz = *tz - cz;
  if (rx > ORA_TILE_RADIUS || rz > ORA_TILE_RADIUS || rx < -ORA_TILE_RADIUS || rz < -ORA_TILE_RADIUS)
    return NULL;
  for (int32_t k = 0; k < g_tiles.n; ++k)
    if (g_tiles.xz[2 * k] == *tx && g_tiles.xz[2 * k + 1] == *tz)
      return g_tiles.atlas + 4 * g_tiles.texels * k;
  return NULL;
}

/* pbr_light.glsl:240-363.  With a baked tile at pos the probes are placed on the
 * tile's grid (:258-275); the texture sampled is `atlas` unless tile_tex is set,
 * in which case it is the tile's own atlas: the deferred lookup passes the baked
 * texture (deferred_tiled.comp:276-283), the probe and mirror closest-hit shaders
 * pass the live atlas (rchit:99-106). */
static f3 sample_probe_t(const ora_config* c, const ora_frame* f, const float* atlas, f3 pos, f3 n,
                         int tile_tex)
{
  int px, py, pz;
  f3 base, step;
  int tx, tz;
  const float* baked = tile_atlas(f, pos, &tx, &tz);
  if (baked) {
    /* pbr_light.glsl:258-275 */
    int tsize = g_tiles.size;
    f3 toff = mk((float)tx * (float)tsize, 0.0f, (float)tz * (float)tsize);
    step = mk((float)tsize / (float)c->probes_x, 1.0f, (float)tsize / (float)c->probes_z);
    f3 pit = sub(pos, toff);
    px = (int)floorf(pit.x / step.x); px = px < 0 ? 0 : (px > c->probes_x - 1 ? c->probes_x - 1 : px);
    py = (int)floorf(pit.y / step.y); py = py < 0 ? 0 : (py > c->probes_y - 1 ? c->probes_y - 1 : py);
    pz = (int)floorf(pit.z / step.z); pz = pz < 0 ? 0 : (pz > c->probes_z - 1 ? c->probes_z - 1 : pz);
    base = add(mul(mk((float)px, (float)py, (float)pz), step), toff);
    if (tile_tex) atlas = baked;
  } else if (f->flags & ORA_FLAG_BAKE) {
    /* pbr_light.glsl:276-293 */
    int tsize = f->bake_tile[2];
    f3 toff = mk((float)f->bake_tile[0] * (float)tsize, 0.0f, (float)f->bake_tile[1] * (float)tsize);
    step = mk((float)tsize / (float)c->probes_x, 1.0f, (float)tsize / (float)c->probes_z);
    f3 pit = sub(pos, toff);
    px = (int)floorf(pit.x / step.x); px = px < 0 ? 0 : (px > c->probes_x - 1 ? c->probes_x - 1 : px);
    py = (int)floorf(pit.y / step.y); py = py < 0 ? 0 : (py > c->probes_y - 1 ? c->probes_y - 1 : py);
    pz = (int)floorf(pit.z / step.z); pz = pz < 0 ? 0 : (pz > c->probes_z - 1 ? c->probes_z - 1 : pz);
    base = add(mul(mk((float)px, (float)py, (float)pz), step), toff);
  } else {
    /* pbr_light.glsl:294-308 */
    step = ld3(c->probe_step);
    f3 off = cam_offset(c, f);
    px = (int)floorf((pos.x - off.x) / step.x); px = px < 0 ? 0 : (px > c->probes_x - 1 ? c->probes_x - 1 : px);
    py = (int)floorf(pos.y / step.y);           py = py < 0 ? 0 : (py > c->probes_y - 1 ? c->probes_y - 1 : py);
    pz = (int)floorf((pos.z - off.z) / step.z); pz = pz < 0 ? 0 : (pz > c->probes_z - 1 ? c->probes_z - 1 : pz);
    base = add(mul(mk((float)px, (float)py, (float)pz), step), off);
  }
  f3 alpha = mk(clampf((pos.x - base.x) / step.x, 0.0f, 1.0f),
                clampf((pos.y - base.y) / step.y, 0.0f, 1.0f),
                clampf((pos.z - base.z) / step.z, 0.0f, 1.0f));
  /* corner order of pbr_light.glsl:318-355 */
  static const int corners[8][3] = {{0, 0, 0}, {1, 1, 1}, {1, 1, 0}, {1, 0, 0},
                                    {0, 1, 1}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}};
  float sumw = 0.0f;
  f3 sumi = mk(0.0f, 0.0f, 0.0f);
  for (int k = 0; k < 8; ++k) {
    int ox = corners[k][0], oy = corners[k][1], oz = corners[k][2];
    f3 ppos = add(base, mul(mk((float)ox, (float)oy, (float)oz), step));
    /* weightProbe, pbr_light.glsl:190-238 */
    f3 tri = mk(ox ? alpha.x : 1.0f - alpha.x, oy ? alpha.y : 1.0f - alpha.y, oz ? alpha.z : 1.0f - alpha.z);
    if (MUTANT(6))
      tri = mk(ox ? 1.0f - alpha.x : alpha.x, oy ? 1.0f - alpha.y : alpha.y, oz ? 1.0f - alpha.z : alpha.z);
    f3 dir = normalize(sub(ppos, pos));
    float x = fmaxf(0.0001f, (dot(dir, n) + 1.0f) * 0.5f);
    float w = x * x + (MUTANT(3) ? 0.0f : 0.2f); /* pow(x, 2) contract: x * x */
    w = fmaxf(0.000001f, w);
    if (w < 0.2f) w = w * (w * w * (1.0f / (0.2f * 0.2f)));
    w = w * (tri.x * tri.y * tri.z);
    int tk = MUTANT(4) ? (k == 2 ? 3 : (k == 3 ? 2 : k)) : k;
    f3 s = sample_single_probe(c, atlas, px + corners[tk][0], py + corners[tk][1], pz + corners[tk][2], n);
    if (!MUTANT(10)) s = mk(sqrtf(s.x), sqrtf(s.y), sqrtf(s.z));
    sumi = add(sumi, muls(s, w));
    sumw = sumw + w;
  }
  f3 net = divs(sumi, sumw);
  if (!MUTANT(10)) net = mul(net, net); /* pow(net, 2) contract: x * x */
  return muls(muls(net, 0.5f), 3.1415926f);
}

void ora_sample_probe(const ora_config* c, const ora_frame* f, const float* atlas, const float* pos,
                      const float* normal, float* out)
{
  st3(out, sample_probe_t(c, f, atlas, ld3(pos), ld3(normal), 0));
}

/* pbr_light.glsl:365-388 */
static f3 calc_indirect_t(const ora_config* c, const ora_frame* f, const float* atlas, f3 N,
                          f3 albedo, float metallic, float rough, f3 pos, f3 view, int tile_tex)
{
  f3 V = normalize(sub(view, pos));
  f3 F0 = add(muls(mk(0.04f, 0.04f, 0.04f), 1.0f - metallic), muls(albedo, metallic));
  float r = MUTANT(7) ? rough : rough * rough; /* pbr_light.glsl:380 passes roughness^2 */
  float p5 = pow5_term(fmaxf(dot(N, V), 0.0f));
  float omr = 1.0f - r;
  f3 kS = mk(F0.x + (fmaxf(omr, F0.x) - F0.x) * p5,
             F0.y + (fmaxf(omr, F0.y) - F0.y) * p5,
             F0.z + (fmaxf(omr, F0.z) - F0.z) * p5);
  f3 kD = muls(sub(mk(1.0f, 1.0f, 1.0f), kS), 1.0f - metallic);
  f3 irr = sample_probe_t(c, f, atlas, pos, N, tile_tex);
  f3 diffuse = mul(irr, albedo);
  return mul(kD, diffuse);
}

static f3 calc_indirect(const ora_config* c, const ora_frame* f, const float* atlas, f3 N,
                        f3 albedo, float metallic, float rough, f3 pos, f3 view)
{
  return calc_indirect_t(c, f, atlas, N, albedo, metallic, rough, pos, view, 0);
}

/* deferred_tiled.comp:264-298 (live atlas, or the baked atlas of the texel's tile):
 * gbuf = n x 12 floats (pos, shadow, normal, metallic, albedo, roughness),
 * out = n x 4 floats (ambient rgb, 0). */
void ora_gi_lookup(const ora_config* c, const ora_frame* f, const float* atlas, const float* gbuf,
                   int64_t n, float* out)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const float* t = gbuf + 12 * i;
    f3 pos = mk(t[0], t[1], t[2]), N = mk(t[4], t[5], t[6]), albedo = mk(t[8], t[9], t[10]);
    float sh = clampf(t[3], 0.3f, 1.0f);
    f3 amb;
    int tx, tz;
    const int baked = tile_atlas(f, pos, &tx, &tz) != NULL; /* deferred_tiled.comp:265-268 */
    if ((f->flags & ORA_FLAG_DDGI) || baked) {
      f3 cam = mk(f->camera_pos[0], f->camera_pos[1], f->camera_pos[2]);
      f3 d = sub(cam, pos);
      float dist = sqrtf(dot(d, d)); /* distance(ubo.cameraPos.xyz, fragWorldPos) */
      if (!baked && dist > (float)c->probes_x * c->probe_step[0] / 2.0f)
        amb = muls(mul(muls(mk(0.3f, 0.3f, 0.3f), f->sky_intensity), albedo), sh);
      else
        amb = calc_indirect_t(c, f, atlas, N, albedo, t[7], t[11], pos, cam, 1);
    } else {
      amb = muls(mul(mk(0.3f, 0.3f, 0.3f), albedo), sh);
    }
    out[4 * i + 0] = amb.x;
    out[4 * i + 1] = amb.y;
    out[4 * i + 2] = amb.z;
    out[4 * i + 3] = 0.0f;
  }
}

/* Radiance of a closest hit seen from `origin`: the probe closest-hit shader
 * (rchit:24-110) or, with mirror set, mirror_reflections.rchit, which differs in
 * the normal transform (mat3(ObjectToWorld) * n, :72) and the indirect-light
 * gate (UBO_DDGI_FLAG, :96).  The shadow ray covers [tmin, tmax]; the world hit
 * position is returned in pos_out (nullable). */
static void shade_surface(const ora_scene* s, const ora_config* c, const ora_frame* f,
                          const float* atlas, const float* origin, const ora_hit* hit, int mirror,
                          float tmin, float tmax, float* out, int32_t* shadowed, uint64_t* counters,
                          f3* pos_out)
{
  if (hit->tri < 0) {
    ora_sky_radiance(f, out);
    if (shadowed) *shadowed = -1;
    return;
  }
  uint32_t k = s->tri_inst[hit->tri];
  const ora_instance* in = &s->inst[k];
  int64_t prim = hit->tri - s->inst_tri_offset[k];
  const ora_mesh* m = &s->meshes[mesh_of(in)];
  const ora_material* mat = &s->mats[mat_of(in)];
  const float* vv[3];
  for (int j = 0; j < 3; ++j) {
    uint32_t idx = s->indices[m->index_offset + 3 * prim + j];
    vv[j] = s->vertices + (int64_t)(m->vertex_offset + idx) * ORA_VERTEX_WORDS;
  }
  /* rchit:40-42 */
  f3 b = mk(1.0f - hit->u - hit->v, hit->u, hit->v);
  f3 pos = add(add(muls(ld3(vv[0]), b.x), muls(ld3(vv[1]), b.y)), muls(ld3(vv[2]), b.z));
  pos = xform_point(in->transform, pos);
  /* rchit:44-66: shadow ray toward the sun, terminate on first hit */
  f3 sdir = normalize(mk(-f->light_dir[0], -f->light_dir[1], -f->light_dir[2]));
  ora_hit sh;
  uint64_t cn = 0, ct = 0;
  trace_one(s, pos, sdir, tmin, tmax, 1, &sh, &cn, &ct);
  if (counters) { counters[0] += cn; counters[1] += ct; }
  if (pos_out) *pos_out = pos;
  int lit = sh.tri < 0; /* rmiss:19 sets shadow = 0 */
  /* rchit:70-74; unpackVertex bindless.glsl:98-109: color = (d0.w, d1.xy), normal = (d1.zw, d2.x) */
  f3 nrm = mk(0.0f, 0.0f, 0.0f), col = mk(0.0f, 0.0f, 0.0f);
  float bw[3] = {b.x, b.y, b.z};
  for (int j = 0; j < 3; ++j) {
    f3 nj = mk(vv[j][6], vv[j][7], vv[j][8]);
    f3 cj = mk(vv[j][3], vv[j][4], vv[j][5]);
    nrm = add(nrm, muls(nj, bw[j]));
    col = add(col, muls(cj, bw[j]));
  }
  nrm = normalize(nrm);
  f3 nw;
  if (mirror) {
    /* mat3(gl_ObjectToWorldEXT) * normal: r[j] = sum_i O2W[j][i] * n[i] */
    const float* t = in->transform;
    nw = mk(t[0] * nrm.x + t[1] * nrm.y + t[2] * nrm.z, t[4] * nrm.x + t[5] * nrm.y + t[6] * nrm.z,
            t[8] * nrm.x + t[9] * nrm.y + t[10] * nrm.z);
  } else {
    /* normal * gl_WorldToObjectEXT: r[j] = sum_i n[i] * W2O[i][j] (row i, column j) */
    const float* wi = s->inv_xform + 12 * (int64_t)k;
    nw = mk(nrm.x * wi[0] + nrm.y * wi[4] + nrm.z * wi[8],
            nrm.x * wi[1] + nrm.y * wi[5] + nrm.z * wi[9],
            nrm.x * wi[2] + nrm.y * wi[6] + nrm.z * wi[10]);
  }
  nrm = normalize(nw);
  col = mk(to_linear1(col.x), to_linear1(col.y), to_linear1(col.z));
  /* getSurfaceDataFromMat, bindless.glsl:181-237, untextured branch */
  f3 albedo = col;
  if (mat->base_col_fac[0] + mat->base_col_fac[1] + mat->base_col_fac[2] > 0.1f)
    albedo = ld3(mat->base_col_fac);
  float metallic = mat->met_rough[0];
  float rough = mat->met_rough[1];
  f3 emissive = muls(ld3(mat->emissive), mat->emissive[3]);
  f3 view = ld3(origin);
  f3 acc = mk(0.0f, 0.0f, 0.0f);
  if (lit) acc = calc_light(f, nrm, albedo, metallic, rough, pos, view);
  if (f->flags & (mirror ? ORA_FLAG_DDGI : ORA_FLAG_DDGI_MULTI))
    acc = add(acc, calc_indirect(c, f, atlas, nrm, albedo, metallic, rough, pos, view));
  acc = add(acc, emissive);
  st3(out, acc);
  if (shadowed) *shadowed = lit ? 0 : 1;
}

void ora_shade_hit(const ora_scene* s, const ora_config* c, const ora_frame* f, const float* atlas,
                   const float* origin, const ora_hit* hit, float* out, int32_t* shadowed,
                   uint64_t* counters)
{
  shade_surface(s, c, f, atlas, origin, hit, 0, c->t_min, c->t_max, out, shadowed, counters, NULL);
}

/* mirror_reflections.rgen:35-82 + .rchit + .rmiss: per pixel of a depth buffer,
 * the world position (as shadow.rgen), the normal from normal4 (xyz of 4 floats,
 * normalized), a closest-hit ray along reflect(normalize(pos - cameraPos), normal)
 * over [0.1, 50], shaded as above; out4 = (irradiance, distance to the hit), a
 * miss gives (sky, 1e6), a sky pixel (depth 1) gives 0.  counters4 (nullable):
 * reflection node visits and triangle tests, shadow node visits and tests. */
void ora_mirror_reflections(const ora_scene* s, const ora_config* c, const ora_frame* f,
                            const float* atlas, const float* depth, const float* normal4, int32_t w,
                            int32_t h, const float* ivp, float* out4, uint64_t* counters4)
{
  uint64_t rn = 0, rt = 0, sn = 0, stt = 0;
  f3 cam = mk(f->camera_pos[0], f->camera_pos[1], f->camera_pos[2]);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : rn, rt, sn, stt)
  for (int64_t i = 0; i < (int64_t)w * h; ++i) {
    float* o = out4 + 4 * i;
    float d = depth[i];
    if (d == 1.0f) { o[0] = o[1] = o[2] = o[3] = 0.0f; continue; }
    int32_t y = (int32_t)(i / w), x = (int32_t)(i - (int64_t)y * w);
    float u = (2.0f * (float)x + 1.0f) / (2.0f * (float)w);
    float v = (2.0f * (float)y + 1.0f) / (2.0f * (float)h);
    float cx = u * 2.0f - 1.0f, cy = v * 2.0f - 1.0f;
    float px = ivp[0] * cx + ivp[4] * cy + ivp[8] * d + ivp[12];
    float py = ivp[1] * cx + ivp[5] * cy + ivp[9] * d + ivp[13];
    float pz = ivp[2] * cx + ivp[6] * cy + ivp[10] * d + ivp[14];
    float pw = ivp[3] * cx + ivp[7] * cy + ivp[11] * d + ivp[15];
    f3 pos = mk(px / pw, py / pw, pz / pw);
    f3 nrm = normalize(ld3(normal4 + 4 * i));
    f3 vv = normalize(sub(pos, cam));
    f3 dir = sub(vv, muls(nrm, 2.0f * dot(nrm, vv))); /* reflect(v, normal) */
    ora_hit hit;
    uint64_t a = 0, b = 0;
    trace_one(s, pos, dir, 0.1f, 50.0f, 0, &hit, &a, &b);
    rn += a; rt += b;
    if (hit.tri < 0) {
      ora_sky_radiance(f, o);
      o[3] = 1e6f;
      continue;
    }
    float origin[3] = {pos.x, pos.y, pos.z};
    uint64_t sc[2] = {0, 0};
    f3 hp;
    shade_surface(s, c, f, atlas, origin, &hit, 1, 0.1f, 50.0f, o, NULL, sc, &hp);
    sn += sc[0]; stt += sc[1];
    f3 dd = sub(hp, pos);
    o[3] = sqrtf(dot(dd, dd)); /* distance(pos, gl_WorldRayOriginEXT) */
  }
  if (counters4) { counters4[0] += rn; counters4[1] += rt; counters4[2] += sn; counters4[3] += stt; }
}

/* ------------------------------------------------------------------------ */
/* RGBA16F storage (Q6, ORA_QUIRK_FP16)                                      */
/* ------------------------------------------------------------------------ */
/* Round to the nearest binary16 value, ties to even, and widen back: the
 * binary16 quantum at binade e is 2^(e - 10) for normal halves and 2^-24 below
 * 2^-14; |x| / quantum is exact in binary64 and rint() rounds it to even.
 * Magnitudes from 65520 (the midpoint above 65504) round to infinity. */
static float f16_round(float x)
{
  uint32_t u;
  memcpy(&u, &x, 4);
  const uint32_t a = u & 0x7fffffffu;
  if (a >= 0x7f800000u) return x;                 /* inf, nan */
  const float sgn = (u >> 31) ? -1.0f : 1.0f;
  if (a >= 0x477ff000u) return sgn * INFINITY;    /* >= 65520 */
  const int e = (int)(a >> 23) - 127;
  const int qe = e < -14 ? -24 : e - 10;
  const double q = ldexp(1.0, qe);
  return sgn * (float)(rint((double)fabsf(x) / q) * q);
}

void ora_f16_round(const float* in, float* out, int64_t n)
{
  for (int64_t i = 0; i < n; ++i) out[i] = f16_round(in[i]);
}

/* ------------------------------------------------------------------------ */
/* Convolution + hysteresis: probe_conv.comp:21-89                           */
/* ------------------------------------------------------------------------ */
void ora_convolve_blend(const ora_config* c, const ora_frame* f, const float* dirs,
                        const float* radiance, float* atlas)
{
  int n = c->rays_per_probe;
  int W = (int)ora_atlas_width(c);
  float td[64 * 3];
  ora_texel_dirs(td);
  int np = c->probes_x * c->probes_z * f->layer_count;
  float h = c->hysteresis;
  int bake = (f->flags & ORA_FLAG_BAKE) != 0;
#pragma omp parallel for schedule(static)
  for (int q = 0; q < np; ++q) {
    int x = q % c->probes_x;
    int z = (q / c->probes_x) % c->probes_z;
    int y = f->layer_begin + q / (c->probes_x * c->probes_z);
    const float* L = radiance + (int64_t)q * 3 * n;
    for (int t = 0; t < 64; ++t) {
      f3 o = ld3(td + 3 * t);
      float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int r = 0; r < n; ++r) {
        float w = fmaxf(0.0f, dot(ld3(dirs + 3 * r), o));
        if (w > 1e-6f) {
          acc[0] = acc[0] + w * L[r];
          acc[1] = acc[1] + w * L[n + r];
          acc[2] = acc[2] + w * L[2 * n + r];
          acc[3] = acc[3] + w;
        }
      }
      int tx = t % 8, ty = t / 8;
      int64_t X = 10 * x + 1 + tx;
      int64_t Y = 10 * ((int64_t)y * c->probes_z + z) + 1 + ty;
      float* o4 = atlas + 4 * (Y * W + X);
      float res[4];
      if (bake) {
        /* probe_conv.comp:67-79 */
        float tw = o4[3] + acc[3];
        for (int a = 0; a < 3; ++a) {
          float v = acc[a] + o4[a] * o4[3];
          if (tw > 1e-6f) v = v / tw;
          res[a] = v;
        }
        res[3] = tw;
      } else {
        /* probe_conv.comp:81-86 */
        if (acc[3] > 1e-6f) { acc[0] = acc[0] / acc[3]; acc[1] = acc[1] / acc[3]; acc[2] = acc[2] / acc[3]; }
        for (int a = 0; a < 4; ++a) res[a] = acc[a] * (1.0f - h) + o4[a] * h;
      }
      if (c->quirks & ORA_QUIRK_FP16)  /* imageStore into the RGBA16F atlas, :89 */
        for (int a = 0; a < 4; ++a) res[a] = f16_round(res[a]);
      memcpy(o4, res, 16);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Translation: IrradianceProbeTranslationPass.cpp:56-186, restated as the   */
/* two vkCmdCopyImage steps (atlas -> temp, then per-layer shifted regions   */
/* temp -> atlas).  Plane height is 10 * probes_z (the reference uses the    */
/* atlas width, equal when probes_x == probes_z).                            */
/* ------------------------------------------------------------------------ */
int ora_translate(const ora_config* c, const float* cam, int32_t* last, float* atlas)
{
  int ix = (int)(cam[0] / c->probe_step[0]);
  int iz = (int)(cam[2] / c->probe_step[2]);
  int dx = ix - last[0], dz = iz - last[1];
  if (dx == 0 && dz == 0) return 0;
  last[0] = ix; last[1] = iz;
  int64_t W = ora_atlas_width(c), H = ora_atlas_height(c);
  float* tmp = (float*)malloc(sizeof(float) * 4 * (size_t)(W * H));
  memcpy(tmp, atlas, sizeof(float) * 4 * (size_t)(W * H));
  const int pps = 10;
  int64_t plane_h = 10 * (int64_t)c->probes_z;
  for (int i = 0; i < c->probes_y; ++i) {
    int64_t py = (int64_t)i * pps * c->probes_z;
    int64_t ew = W, eh = plane_h, sx = 0, sy = py, ox = 0, oy = py;
    if (dx > 0) { ew = W - pps; sx = pps; ox = 0; }
    else if (dx < 0) { ew = W - pps; sx = 0; ox = pps; }
    if (dz > 0) { eh = plane_h - pps; sy += pps; }
    else if (dz < 0) { eh = plane_h - pps; oy += pps; }
    for (int64_t yy = 0; yy < eh; ++yy)
      memcpy(atlas + 4 * ((oy + yy) * W + ox), tmp + 4 * ((sy + yy) * W + sx), sizeof(float) * 4 * (size_t)ew);
  }
  free(tmp);
  return 1;
}

/* ------------------------------------------------------------------------ */
/* One update (IrradianceProbeRT then IrradianceProbeConvolve).              */
/* ------------------------------------------------------------------------ */
/* wall-clock ms of the last ora_update's two stages: [0] trace + shadow + shade,
 * [1] convolve + blend (the CPU baseline's per-stage figures) */
static double g_stage_ms[2];

static double now_ms(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

void ora_last_update_stage_ms(double* out2)
{
  out2[0] = g_stage_ms[0];
  out2[1] = g_stage_ms[1];
}

void ora_update(const ora_scene* s, const ora_config* c, const ora_frame* f, float* atlas,
                ora_hit* hits, uint8_t* shadowed, float* radiance, uint64_t* counters)
{
  const double t0 = now_ms();
  int n = c->rays_per_probe;
  float* base = (float*)malloc(sizeof(float) * 3 * (size_t)n);
  float* dirs = (float*)malloc(sizeof(float) * 3 * (size_t)n);
  const int fp16 = (c->quirks & ORA_QUIRK_FP16) != 0;
  ora_fibonacci_dirs(n, base);
  ora_rotate_dirs(n, f->rotation, base, dirs);
  int64_t np = (int64_t)c->probes_x * c->probes_z * f->layer_count;
  float* rad = radiance ? radiance : (float*)malloc(sizeof(float) * 3 * (size_t)(np * n));
  uint64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : c0, c1, c2, c3)
  for (int64_t q = 0; q < np; ++q) {
    int x = (int)(q % c->probes_x);
    int z = (int)((q / c->probes_x) % c->probes_z);
    int y = f->layer_begin + (int)(q / ((int64_t)c->probes_x * c->probes_z));
    float org[3];
    ora_probe_origin(c, f, x, y, z, org);
    for (int r = 0; r < n; ++r) {
      ora_hit h;
      uint64_t pc[2] = {0, 0};
      trace_one(s, ld3(org), ld3(dirs + 3 * r), c->t_min, c->t_max, 0, &h, &pc[0], &pc[1]);
      c0 += pc[0]; c1 += pc[1];
      float L[3];
      int32_t sh = -1;
      uint64_t sc[2] = {0, 0};
      ora_shade_hit(s, c, f, atlas, org, &h, L, &sh, sc);
      c2 += sc[0]; c3 += sc[1];
      float* R = rad + q * 3 * n;
      if (fp16) { L[0] = f16_round(L[0]); L[1] = f16_round(L[1]); L[2] = f16_round(L[2]); } /* rgen:109 */
      R[r] = L[0]; R[n + r] = L[1]; R[2 * n + r] = L[2];
      if (hits) hits[q * n + r] = h;
      if (shadowed) shadowed[q * n + r] = (uint8_t)(sh < 0 ? 255 : sh);
    }
  }
  const double t1 = now_ms();
  if (fp16) /* the convolution reads the directions back from the RGBA16F image, rgen:108 */
    for (int k = 0; k < 3 * n; ++k) base[k] = f16_round(dirs[k]);
  ora_convolve_blend(c, f, fp16 ? base : dirs, rad, atlas);
  g_stage_ms[0] = t1 - t0;
  g_stage_ms[1] = now_ms() - t1;
  if (counters) { counters[0] += c0; counters[1] += c1; counters[2] += c2; counters[3] += c3; }
  if (!radiance) free(rad);
  free(base);
  free(dirs);
}

int ora_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
