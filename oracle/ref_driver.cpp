/*
 * ref_driver.cpp -- builds the reference's own host path (include/geomfunc.h, the CPU twin of the
 * rendering kernel) into oracle/_ref/libref.so for cross-checking the restatement in oracle.cpp
 * and for timing the reference CPU path.  TEST INFRASTRUCTURE ONLY.
 *
 * Nothing from the reference is copied: the reference headers are included from the reference
 * tree at build time (see oracle/Makefile).  The reference configures itself with compile-time
 * macros (geom.h:32-35: MAX_DEPTH, ACCELSTR); this driver is compiled once per
 * (REF_ACCEL, REF_DEPTH) pair and sets those two macros the same way a reference build would.
 * Each build lives in its own namespace and exports C wrappers suffixed _a<accel>_d<depth>.
 *
 * The per-pixel loop below follows RadianceGPU (rendering_kernel.cl:1100-1149) with the camera
 * ray of DrawBox (smallptGPU.cpp:668-689) and the GPU seed slot 2*(y*W+x).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef REF_ACCEL
#error REF_ACCEL must be defined
#endif
#ifndef REF_DEPTH
#error REF_DEPTH must be defined
#endif

namespace REF_NS {
#include "geom.h"
#undef MAX_DEPTH
#define MAX_DEPTH REF_DEPTH
#undef ACCELSTR
#define ACCELSTR REF_ACCEL
#include "camera.h"
#include "BVHNodeGPU.h"
#include "KDNodeGPU.h"
#include "geomfunc.h"
}

#define REF_CAT2(a, b) a##b
#define REF_CAT(a, b) REF_CAT2(a, b)
#define REF_SUFFIX REF_CAT(REF_CAT(_a, REF_ACCEL), REF_CAT(_d, REF_DEPTH))
#define REF_FN(name) REF_CAT(name, REF_SUFFIX)

using namespace REF_NS;

static int ref_intersect_dispatch(const Shape *shapes, unsigned int n, const void *a, const void *b,
                                  int szkng, int szkn, const Ray *ray, float *t, unsigned int *id) {
#if REF_ACCEL == 0
    (void)a; (void)b; (void)szkng; (void)szkn;
    return Intersect(shapes, n, ray, t, id);
#elif REF_ACCEL == 1
    (void)szkng; (void)szkn;
    return Intersect(shapes, n, (BVHNodeGPU *)a, (BVHNodeGPU *)b, ray, t, id);
#else
    return Intersect(shapes, n, (KDNodeGPU *)a, (int *)b, szkng, szkn, ray, t, id);
#endif
}

static void ref_path(const Shape *shapes, unsigned int n, unsigned int lightCnt, const void *a,
                     const void *b, int szkng, int szkn, const Ray *ray, unsigned int *s0,
                     unsigned int *s1, Vec *result) {
#if REF_ACCEL == 0
    (void)a; (void)b; (void)szkng; (void)szkn;
    RadiancePathTracing(shapes, n, lightCnt, ray, s0, s1, result);
#elif REF_ACCEL == 1
    (void)szkng; (void)szkn;
    RadiancePathTracing(shapes, n, lightCnt, (BVHNodeGPU *)a, (BVHNodeGPU *)b, ray, s0, s1, result);
#else
    RadiancePathTracing(shapes, n, lightCnt, (KDNodeGPU *)a, (int *)b, szkng, szkn, ray, s0, s1,
                        result);
#endif
}

extern "C" {

float REF_FN(ref_get_random)(unsigned int *s0, unsigned int *s1) { return GetRandom(s0, s1); }

float REF_FN(ref_sphere_intersect)(const void *sph, const void *ray) {
    return SphereIntersect((const Sphere *)sph, (const Ray *)ray);
}

float REF_FN(ref_triangle_intersect)(const void *tri, const void *ray) {
    return TriangleIntersect((const Triangle *)tri, (const Ray *)ray);
}

int REF_FN(ref_bound_test)(const void *ray, const void *bound) {
    return intersection_bound_test(*(const Ray *)ray, *(const Bound *)bound) ? 1 : 0;
}

int REF_FN(ref_intersect)(const void *shapes, unsigned int n, const void *a, const void *b,
                          int szkng, int szkn, const void *ray, float *t, unsigned int *id) {
    return ref_intersect_dispatch((const Shape *)shapes, n, a, b, szkng, szkn, (const Ray *)ray, t,
                                  id);
}

int REF_FN(ref_intersect_p)(const void *shapes, unsigned int n, const void *ray, float maxt) {
    return IntersectP((const Shape *)shapes, n, (const Ray *)ray, maxt);
}

void REF_FN(ref_radiance_path)(const void *shapes, unsigned int n, unsigned int lightCnt,
                               const void *a, const void *b, int szkng, int szkn, const void *ray,
                               unsigned int *s0, unsigned int *s1, float *result3) {
    Vec r;
    ref_path((const Shape *)shapes, n, lightCnt, a, b, szkng, szkn, (const Ray *)ray, s0, s1, &r);
    result3[0] = r.x; result3[1] = r.y; result3[2] = r.z;
}

/* `spp` samples for each pixel in pix_list (or all pixels), RadianceGPU order per pixel. */
void REF_FN(ref_render)(float *colors, unsigned int *seeds, const void *camera_v,
                        const void *shapes_v, unsigned int n, unsigned int lightCnt, const void *a,
                        const void *b, int szkng, int szkn, int width, int height,
                        int currentSample, int spp, const int *pix_list, int npix, int nthreads) {
    const Camera *camera = (const Camera *)camera_v;
    const Shape *shapes = (const Shape *)shapes_v;
    const float invWidth = 1.f / width;
    const float invHeight = 1.f / height;
    const int total = pix_list ? npix : width * height;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (int q = 0; q < total; q++) {
        const int gid = pix_list ? pix_list[q] : q;
        const int x = gid % width, y = gid / width;
        const int i = (height - y - 1) * width + x;
        unsigned int s0 = seeds[2 * gid], s1 = seeds[2 * gid + 1];
        for (int k = 0; k < spp; k++) {
            const int cs = currentSample + k;
            const float r1 = GetRandom(&s0, &s1) - .5f;
            const float r2 = GetRandom(&s0, &s1) - .5f;
            const float kcx = (x + r1) * invWidth - .5f;
            const float kcy = (y + r2) * invHeight - .5f;
            Vec rdir;
            vinit(rdir, camera->x.x * kcx + camera->y.x * kcy + camera->dir.x,
                  camera->x.y * kcx + camera->y.y * kcy + camera->dir.y,
                  camera->x.z * kcx + camera->y.z * kcy + camera->dir.z);
            Vec rorig;
            vsmul(rorig, 0.1f, rdir);
            vadd(rorig, rorig, camera->orig)
            vnorm(rdir);
            const Ray ray = {rorig, rdir};
            Vec r;
            ref_path(shapes, n, lightCnt, a, b, szkng, szkn, &ray, &s0, &s1, &r);
            float *c = colors + 3 * i;
            if (cs == 0) {
                c[0] = r.x; c[1] = r.y; c[2] = r.z;
            } else {
                const float k1 = cs;
                const float k2 = 1.f / (k1 + 1.f);
                c[0] = (c[0] * k1 + r.x) * k2;
                c[1] = (c[1] * k1 + r.y) * k2;
                c[2] = (c[2] * k1 + r.z) * k2;
            }
        }
        seeds[2 * gid] = s0;
        seeds[2 * gid + 1] = s1;
    }
}

} /* extern "C" */
