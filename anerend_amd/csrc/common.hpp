// common.hpp -- shared device helpers and internal records of libane_ddgi.
//
// Floating-point contract (shared with the CPU oracle so that hit ids, shadow
// decisions, BVH topology and translation are bit-exact):
//   * every float op is one IEEE binary32 op evaluated left to right; the
//     library is compiled with -ffp-contract=off, correctly rounded divide and
//     sqrt, and denormals preserved;
//   * normalize(v) = v / sqrt(dot(v, v)), dot = x*x + y*y + z*z;
//   * GLSL pow(x, 2) is x * x; pow(x, 5.0) and pow(x, 2.4) are powf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ane_ddgi.h"

#define ANE_HD __host__ __device__ __forceinline__

namespace ane {

struct f3 {
  float x, y, z;
};

ANE_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
ANE_HD f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
ANE_HD f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
ANE_HD f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
ANE_HD f3 muls(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
ANE_HD f3 divs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
ANE_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ANE_HD f3 normalize(f3 v) { float l = sqrtf(dot(v, v)); return divs(v, l); }
ANE_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
ANE_HD f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
ANE_HD float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// mat3(m) * v, m a column-major mat4 (16 floats)
ANE_HD f3 rot3(const float* m, f3 v)
{
  return mk(m[0] * v.x + m[4] * v.y + m[8] * v.z,
            m[1] * v.x + m[5] * v.y + m[9] * v.z,
            m[2] * v.x + m[6] * v.y + m[10] * v.z);
}

// 3x4 row-major affine transform of a point (rchit:42 gl_ObjectToWorldEXT * vec4(p, 1))
ANE_HD f3 xform_point(const float* m, f3 p)
{
  return mk(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
            m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}

// Rotation and frame constants passed by value to kernels.
struct Mat3 {
  float m[16];
};

struct FrameK {
  float cam[4];
  float light_dir[4];
  float sky[3];        // toLinear(skyIntensity * skyColor), rmiss:20
  float sun_intensity;
  uint32_t flags;
  int32_t layer_begin, layer_count;
  int32_t bake_tile[4];
  float rot[16];
  float sun_l[3];      // normalize(-light_dir): direction to the sun (pbr_light.glsl:103)
  float sky_intensity; // ubo.skyIntensity as given (deferred_tiled.comp:272)
  // baked tile atlases (worldPosToTileInfo, bindless.glsl:291-308): the window of
  // (2 * ANE_TILE_RADIUS + 1)^2 tiles around the camera's tile, as the renderer
  // fills GPUTileInfo each frame (VulkanRenderer.cpp:3698-3738)
  int32_t tile_cam[2];   // floor(camera x, z / tile_size)
  int32_t tile_size;     // TILE_SIZE, metres
  int32_t tile_count;    // window entries that hold a baked atlas (0: no lookup)
  int32_t tile_win[25];  // baked atlas of tile cam + (i % 5 - 2, i / 5 - 2), -1 none
};

#define ANE_TILE_RADIUS 2  // MAX_PAGE_TILE_RADIUS, VulkanRenderer.h:267

struct GridK {
  int32_t nx, ny, nz;
  float step[3];
  int32_t rays;
  float hysteresis;
  float t_min, t_max;
  int32_t atlas_w, atlas_h;
  // texture bounds of the probe tiles (sampleSingleProbe, pbr_light.glsl:177-183):
  // (start, end - start) of tile column 0..nx and of tile row (y, z) at
  // y * (nz + 1) + z for y 0..ny, z 0..nz; device memory
  const float2* tile_u;
  const float2* tile_v;
  // baked tile atlases: [tiles][atlas_h][atlas_w] RGBA16F texels (4 x binary16)
  const uint2* tile_atlas;
};

// Device view of a scene (all pointers device memory).
struct SceneK {
  const float* vertices;        // [nvert][24]
  const uint32_t* indices;
  const ane_mesh* meshes;
  const ane_material* mats;
  const ane_instance* inst;
  const float* inv_xform;       // [ninst][12] world-to-object 3x4
  const uint32_t* tri_inst;     // [ntri] instance of triangle id
  const float4* tri_rec;        // [ntri][4] of triangle id: (object-space p_j, vertex index j) j = 0..2, (instance)
  const uint32_t* inst_tri_off; // [ninst + 1]
  const float4* nodes;          // [ntri - 1][4] traversal nodes
  const uint4* qnodes;          // [ntri - 1][2] quantized traversal nodes, or null (use nodes)
  const float4* tris;           // [ntri][3] sorted world triangles (v0.w = id bits)
  int32_t ntri;
  int32_t ninst;
};

// Traversal node (Aila-Laine child-box layout), 64 bytes:
//   a = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   b = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   c = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   d = (left, right, parent, 0) as int bits; child >= 0 internal, < 0 leaf ~pos
struct NodeBuild {
  float4 a, b, c;
  int4 d;
};

// Quantized traversal node, 32 bytes (2 x uint4), for scenes whose 64-byte nodes
// outgrow the caches:
//   q0 = (o.x, o.y, o.z as bits, (ex+128) | (ey+128) << 8 | (ez+128) << 16 | leaf0 << 24 | leaf1 << 25)
//   q1 = (12 bound bytes c0.lo.xyz, c0.hi.xyz, c1.lo.xyz, c1.hi.xyz, lowest byte first; split g)
// o is the node box's lower corner; per axis 2^e is the smallest step (e >= -126)
// with 254 * 2^e >= the node box extent.  Lower bounds are floor((v - o) / 2^e),
// upper bounds ceil((v - o) / 2^e), exact in binary64; a bound decodes as
// o + q * 2^e in binary32 (exact product, one rounding), which is monotone, so a
// decoded box contains its exact child box.  The children are g and g + 1 (the
// Karras split), leaves (~position) where the leaf flag is set.
ANE_HD float qnode_step(uint32_t ebyte) { return __builtin_bit_cast(float, (ebyte - 128u + 127u) << 23); }
ANE_HD float qnode_bound(float o, uint32_t q, float step) { return o + (float)q * step; }

ANE_HD uint32_t mesh_of(uint32_t custom) { return (custom >> 10) & 0x3FFFu; }
ANE_HD uint32_t mat_of(uint32_t custom) { return custom & 0x3FFu; }

// toLinear, bindless.glsl:240-247
ANE_HD float to_linear1(float c)
{
  return (c < 0.04045f) ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

// Ray precomputation for the slab and watertight tests.
struct RayPre {
  f3 o, d, inv;
  int kx, ky, kz;
  float sx, sy, sz;
};

ANE_HD void ray_setup(RayPre& r, f3 o, f3 d)
{
  r.o = o;
  r.d = d;
  float dx = fabsf(d.x) < 1e-30f ? copysignf(1e-30f, d.x) : d.x;
  float dy = fabsf(d.y) < 1e-30f ? copysignf(1e-30f, d.y) : d.y;
  float dz = fabsf(d.z) < 1e-30f ? copysignf(1e-30f, d.z) : d.z;
  r.inv = mk(1.0f / dx, 1.0f / dy, 1.0f / dz);
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz + 1;
  if (kx == 3) kx = 0;
  int ky = kx + 1;
  if (ky == 3) ky = 0;
  if (comp(d, kz) < 0.0f) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  float dkz = comp(d, kz);
  r.sx = comp(d, kx) / dkz;
  r.sy = comp(d, ky) / dkz;
  r.sz = 1.0f / dkz;
}

#define ANE_ROBUST 1.0000004f  // 1 + 2*gamma(3) rounded up (Ize 2013)

// Slab test of one box (lo/hi); returns hit and tnear (clamped to tmin).
ANE_HD bool box_test(const RayPre& r, float lx, float hx, float ly, float hy, float lz, float hz,
                     float tmin, float tmax, float& tnear)
{
  float tx0 = (lx - r.o.x) * r.inv.x, tx1 = (hx - r.o.x) * r.inv.x;
  float ty0 = (ly - r.o.y) * r.inv.y, ty1 = (hy - r.o.y) * r.inv.y;
  float tz0 = (lz - r.o.z) * r.inv.z, tz1 = (hz - r.o.z) * r.inv.z;
  float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
  float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  tnear = tn;
  return tn <= tf * ANE_ROBUST;
}

// Watertight ray/triangle test (Woop, Benthin, Wald 2013), no culling.
ANE_HD bool tri_test(const RayPre& r, f3 v0, f3 v1, f3 v2, float tmin, float tmax, float& t_out,
                     float& u_out, float& v_out)
{
  f3 A = sub(v0, r.o), B = sub(v1, r.o), C = sub(v2, r.o);
  float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  float Ax = comp(A, r.kx) - r.sx * Akz, Ay = comp(A, r.ky) - r.sy * Akz;
  float Bx = comp(B, r.kx) - r.sx * Bkz, By = comp(B, r.ky) - r.sy * Bkz;
  float Cx = comp(C, r.kx) - r.sx * Ckz, Cy = comp(C, r.ky) - r.sy * Ckz;
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
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
  float T = U * Az + V * Bz + W * Cz;
  const float rcp = 1.0f / det;  // one correctly rounded reciprocal, then products
  float t = T * rcp;
  if (!(t >= tmin && t <= tmax)) return false;
  t_out = t;
  u_out = V * rcp;
  v_out = W * rcp;
  return true;
}

// Probe camera offset, rgen:75-81 / pbr_light.glsl:295-301 (int() truncates toward zero)
ANE_HD f3 cam_offset(const GridK& g, const float* cam)
{
  int cx = (int)(cam[0] / g.step[0]);
  int cz = (int)(cam[2] / g.step[2]);
  float ox = (float)cx * g.step[0] - g.step[0] * (float)g.nx / 2.0f;
  float oz = (float)cz * g.step[2] - g.step[2] * (float)g.nz / 2.0f;
  return mk(ox, 0.0f, oz);
}

// Probe origin, rgen:60-84
ANE_HD f3 probe_origin(const GridK& g, const FrameK& f, int px, int py, int pz)
{
  f3 p = mk((float)px, (float)py, (float)pz);
  if (f.flags & ANE_FLAG_BAKE) {
    int ts = f.bake_tile[2];
    f3 scale = mk((float)ts / (float)g.nx, 1.0f, (float)ts / (float)g.nz);
    f3 toff = mk((float)f.bake_tile[0] * (float)ts, 0.0f, (float)f.bake_tile[1] * (float)ts);
    return add(mul(p, scale), toff);
  }
  return add(mul(p, mk(g.step[0], g.step[1], g.step[2])), cam_offset(g, f.cam));
}

ANE_HD float sign_not_zero(float k) { return (k >= 0.0f) ? 1.0f : -1.0f; }

// octEncode, octahedron_helpers.glsl:10-20
ANE_HD void oct_encode(f3 vin, float& rx, float& ry)
{
  f3 v = mk(vin.x, vin.z, vin.y);
  float l1 = fabsf(v.x) + fabsf(v.y) + fabsf(v.z);
  float inv = 1.0f / l1;
  float x = v.x * inv, y = v.y * inv;
  if (v.z < 0.0f) {
    float nx = (1.0f - fabsf(y)) * sign_not_zero(x);
    float ny = (1.0f - fabsf(x)) * sign_not_zero(y);
    x = nx;
    y = ny;
  }
  rx = x;
  ry = y;
}

// octDecode, octahedron_helpers.glsl:25-34
ANE_HD f3 oct_decode(float ox, float oy)
{
  f3 v = mk(ox, oy, 1.0f - fabsf(ox) - fabsf(oy));
  if (v.z < 0.0f) {
    float nx = (1.0f - fabsf(v.y)) * sign_not_zero(v.x);
    float ny = (1.0f - fabsf(v.x)) * sign_not_zero(v.y);
    v.x = nx;
    v.y = ny;
  }
  return normalize(mk(v.x, v.z, v.y));
}

// Texel t = ty*8 + tx of the 8x8 octahedral tile: probe_conv.comp:43-46, helpers.glsl:1-4
ANE_HD f3 texel_dir(int t)
{
  int tx = t & 7, ty = t >> 3;
  float u = (2.0f * (float)tx + 1.0f) / (2.0f * 8.0f);
  float v = (2.0f * (float)ty + 1.0f) / (2.0f * 8.0f);
  u = u * 2.0f - 1.0f;
  v = v * 2.0f - 1.0f;
  return oct_decode(u, v);
}

}  // namespace ane

#define ANE_CHECK_HIP(expr)                                               \
  do {                                                                    \
    hipError_t _e = (expr);                                               \
    if (_e != hipSuccess) return ane::fail_hip(_e, #expr, __FILE__, __LINE__); \
  } while (0)

namespace ane {
int fail(int status, const char* fmt, ...);
int fail_hip(hipError_t e, const char* expr, const char* file, int line);
}  // namespace ane
