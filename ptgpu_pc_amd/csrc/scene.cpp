// scene.cpp -- scene ingest and image output around the path (host code, no GPU work).
//
// ptgpu_scene_read reads the reference's text scene format (displayfunc.cpp:144-505, ReadTxt):
//   gamma g  sky1 r g b  sky2 r g b          (optional)
//   camera ox oy oz  tx ty tz
//   materials n, then n lines "matte kd.xyz sigma" / "plastic kd.xyz ks.xyz roughness"
//   objects n, then n object lines:
//     sphere   radius  cx cy cz  emission.xyz  colour.xyz  material
//     triangle p1.xyz  p2.xyz  p3.xyz  emission.xyz  colour.xyz  material
//     model    path  position.xyz  scale.xyz  emission.xyz  colour.xyz  material
// Lines after the n-th object are ignored, as the reference ignores them.  A model expands into one
// triangle per face of its OBJ file, scaled then translated component-wise; OBJ numbers are parsed
// the way the reference's bundled reader does it (tinyobj_loader_c.h:267-391: integer digits by
// m = 10 m + d, fractional digits by m += d * 0.1^k with 0.1^k formed by repeated products, in
// double, then rounded to float), so the vertices are bit-identical.  Material ids 0/1/2 map to
// DIFF/SPEC/REFR and the others to DIFF (displayfunc.cpp:114-129, 458-502).  Areas: 4 pi r^2 for
// spheres, |e1 x e2| / 2 for triangles, in the reference's float operation order.
//
// ptgpu_write_ppm writes the reference's PPM (smallptGPU.cpp:1238-1258): "P3", rows from the top of
// the image (the last row of the pixel buffer) down, "r g b " per pixel from the packed bytes.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

using namespace ptgpu;

namespace {

struct Tokens {
    std::vector<std::string> t;
    size_t i = 0;
    bool next(std::string *out) {
        if (i >= t.size()) return false;
        *out = t[i++];
        return true;
    }
    bool peek_is(const char *s) const { return i < t.size() && t[i] == s; }
};

bool read_file(const std::string &path, std::string *out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char buf[65536];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) out->append(buf, k);
    std::fclose(f);
    return true;
}

void split_ws(const std::string &s, std::vector<std::string> *out) {
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
        size_t j = i;
        while (j < s.size() && !std::isspace((unsigned char)s[j])) j++;
        if (j > i) out->push_back(s.substr(i, j - i));
        i = j;
    }
}

// a scene-file number: the C library's decimal-to-float conversion (what fscanf("%f") does)
bool scene_float(Tokens &tk, float *v) {
    std::string s;
    if (!tk.next(&s)) return false;
    char *end = nullptr;
    errno = 0;
    *v = std::strtof(s.c_str(), &end);
    return end && *end == '\0' && end != s.c_str();
}

bool scene_uint(Tokens &tk, long long *v) {
    std::string s;
    if (!tk.next(&s)) return false;
    char *end = nullptr;
    *v = std::strtoll(s.c_str(), &end, 10);
    return end && *end == '\0' && end != s.c_str();
}

bool scene_vec(Tokens &tk, Vec *v) { return scene_float(tk, &v->x) && scene_float(tk, &v->y) && scene_float(tk, &v->z); }

// an OBJ number as the bundled reader forms it; a negative exponent, which the reader turns into a
// shift by a negative count (undefined), scales by 2^e here
float obj_float(const char *s, const char *e) {
    double m = 0.0;
    int exponent = 0;
    char sign = '+', esign = '+';
    const char *c = s;
    if (c >= e) return 0.f;
    if (*c == '+' || *c == '-') sign = *c++;
    else if (!(*c >= '0' && *c <= '9')) return 0.f;
    int read = 0;
    while (c < e && *c >= '0' && *c <= '9') {
        m *= 10;
        m += (int)(*c - '0');
        c++;
        read++;
    }
    if (read == 0) return 0.f;
    if (c < e && *c == '.') {
        c++;
        read = 1;
        while (c < e && *c >= '0' && *c <= '9') {
            double frac = 1.0;
            for (int f = 0; f < read; f++) frac *= 0.1;
            m += (int)(*c - '0') * frac;
            read++;
            c++;
        }
    }
    if (c < e && (*c == 'e' || *c == 'E')) {
        c++;
        if (c < e && (*c == '+' || *c == '-')) esign = *c++;
        while (c < e && *c >= '0' && *c <= '9') {
            exponent = exponent * 10 + (int)(*c - '0');
            c++;
        }
        if (esign == '-') exponent = -exponent;
    }
    double a = 1.0;
    for (int i = 0; i < exponent; i++) a *= 5.0;
    const double scale = exponent >= 0 && exponent < 31 ? (double)(1 << exponent) : std::ldexp(1.0, exponent);
    const double val = (sign == '+' ? 1 : -1) * (m * a) * scale;
    return (float)val;
}

// vertices and triangle faces (0-based vertex indices) of an OBJ file; faces must be triangles
bool read_obj(const std::string &path, std::vector<float> *verts, std::vector<int> *faces, std::string *why) {
    std::string text;
    if (!read_file(path, &text)) {
        *why = "cannot open model " + path;
        return false;
    }
    size_t i = 0;
    while (i < text.size()) {
        size_t j = text.find('\n', i);
        if (j == std::string::npos) j = text.size();
        std::string line = text.substr(i, j - i);
        i = j + 1;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::vector<std::string> tok;
        split_ws(line, &tok);
        if (tok.empty()) continue;
        if (tok[0] == "v" && tok.size() >= 4) {
            for (int k = 1; k <= 3; k++)
                verts->push_back(obj_float(tok[k].c_str(), tok[k].c_str() + tok[k].size()));
        } else if (tok[0] == "f") {
            if (tok.size() != 4) {
                *why = "model " + path + ": only triangle faces are supported";
                return false;
            }
            const int nv = (int)(verts->size() / 3);
            for (int k = 1; k <= 3; k++) {
                const int idx = std::atoi(tok[k].c_str());   // "v", "v/vt", "v//vn", "v/vt/vn"
                const int v = idx > 0 ? idx - 1 : nv + idx;
                if (idx == 0 || v < 0 || v >= nv) {
                    *why = "model " + path + ": vertex index out of range";
                    return false;
                }
                faces->push_back(v);
            }
        }
    }
    return true;
}

Refl refl_of(long long mat) { return mat == 1 ? SPEC : mat == 2 ? REFR : DIFF; }

float tri_area(const Vec &p1, const Vec &p2, const Vec &p3) {
    const Vec e1 = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
    const Vec e2 = {p3.x - p1.x, p3.y - p1.y, p3.z - p1.z};
    const Vec n = {e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
    return std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z) * 0.5f;
}

bool is_zero(const Vec &v) { return v.x == 0.f && v.y == 0.f && v.z == 0.f; }

}  // namespace

extern "C" {

int ptgpu_scene_read(const char *path, const char *base_dir, int width, int height, Shape **h_shapes,
                     unsigned int *shapeCnt, unsigned int *lightCnt, Camera *camera) {
    PT_CHECK(path && h_shapes && shapeCnt && lightCnt && camera, PTGPU_ERR_INVALID, "null argument");
    *h_shapes = nullptr;
    std::string text;
    PT_CHECK(read_file(path, &text), PTGPU_ERR_INVALID, "cannot open scene %s", path);
    Tokens tk;
    split_ws(text, &tk.t);
    float f;
    Vec v;
    std::string s;
    if (tk.peek_is("gamma")) {   // gamma g sky1 r g b sky2 r g b (optional)
        tk.i++;
        PT_CHECK(scene_float(tk, &f) && tk.next(&s) && s == "sky1" && scene_vec(tk, &v) && tk.next(&s) &&
                     s == "sky2" && scene_vec(tk, &v),
                 PTGPU_ERR_INVALID, "%s: malformed gamma / sky line", path);
    }
    std::memset(camera, 0, sizeof(*camera));
    PT_CHECK(tk.next(&s) && s == "camera" && scene_vec(tk, &camera->orig) && scene_vec(tk, &camera->target),
             PTGPU_ERR_INVALID, "%s: failed to read the 6 camera parameters", path);
    camera->width = (float)width;
    camera->height = (float)height;
    long long nmat = 0, nobj = 0;
    PT_CHECK(tk.next(&s) && s == "materials" && scene_uint(tk, &nmat) && nmat >= 0, PTGPU_ERR_INVALID,
             "%s: failed to read the number of materials", path);
    for (long long m = 0; m < nmat; m++) {   // parsed and dropped: materials never reach the kernels
        PT_CHECK(tk.next(&s) && (s == "matte" || s == "plastic"), PTGPU_ERR_INVALID, "%s: material %lld type", path,
                 m);
        const int nf = s == "matte" ? 4 : 7;
        for (int k = 0; k < nf; k++)
            PT_CHECK(scene_float(tk, &f), PTGPU_ERR_INVALID, "%s: material %lld values", path, m);
    }
    PT_CHECK(tk.next(&s) && s == "objects" && scene_uint(tk, &nobj) && nobj >= 0, PTGPU_ERR_INVALID,
             "%s: failed to read the object count", path);
    std::string base = base_dir && *base_dir ? std::string(base_dir) : std::string(".");
    if (base.back() != '/') base += '/';
    std::vector<Shape> out;
    unsigned lights = 0;
    auto shape0 = [](Shape *sh) { std::memset(sh, 0, sizeof(*sh)); };
    for (long long o = 0; o < nobj; o++) {
        PT_CHECK(tk.next(&s), PTGPU_ERR_INVALID, "%s: object %lld missing", path, o);
        long long mat = 0;
        Vec e, c;
        if (s == "sphere") {
            Shape sh;
            shape0(&sh);
            sh.type = SPHERE;
            PT_CHECK(scene_float(tk, &sh.s.rad) && scene_vec(tk, &sh.s.p) && scene_vec(tk, &e) && scene_vec(tk, &c) &&
                         scene_uint(tk, &mat) && mat >= 0 && mat <= 5,
                     PTGPU_ERR_INVALID, "%s: failed to read sphere #%lld", path, o);
            sh.e = e;
            sh.c = c;
            sh.refl = refl_of(mat);
            sh.area = 4 * 3.14159265358979323846f * sh.s.rad * sh.s.rad;
            if (!is_zero(e)) lights++;
            out.push_back(sh);
        } else if (s == "triangle") {
            Shape sh;
            shape0(&sh);
            sh.type = TRIANGLE;
            PT_CHECK(scene_vec(tk, &sh.t.p1) && scene_vec(tk, &sh.t.p2) && scene_vec(tk, &sh.t.p3) && scene_vec(tk, &e) &&
                         scene_vec(tk, &c) && scene_uint(tk, &mat) && mat >= 0 && mat <= 5,
                     PTGPU_ERR_INVALID, "%s: failed to read triangle #%lld", path, o);
            sh.e = e;
            sh.c = c;
            sh.refl = refl_of(mat);
            sh.area = tri_area(sh.t.p1, sh.t.p2, sh.t.p3);
            if (!is_zero(e)) lights++;
            out.push_back(sh);
        } else if (s == "model") {
            std::string mpath;
            Vec pos, scl;
            PT_CHECK(tk.next(&mpath) && scene_vec(tk, &pos) && scene_vec(tk, &scl) && scene_vec(tk, &e) &&
                         scene_vec(tk, &c) && scene_uint(tk, &mat) && mat >= 0 && mat <= 5,
                     PTGPU_ERR_INVALID, "%s: failed to read model #%lld", path, o);
            std::vector<float> verts;
            std::vector<int> faces;
            std::string why;
            const std::string full = mpath.size() && mpath[0] == '/' ? mpath : base + mpath;
            PT_CHECK(read_obj(full, &verts, &faces, &why), PTGPU_ERR_INVALID, "%s", why.c_str());
            for (size_t t = 0; t + 2 < faces.size(); t += 3) {
                Shape sh;
                shape0(&sh);
                sh.type = TRIANGLE;
                Vec *p[3] = {&sh.t.p1, &sh.t.p2, &sh.t.p3};
                for (int k = 0; k < 3; k++) {
                    const float *q = &verts[(size_t)faces[t + k] * 3];
                    Vec w = {q[0] * scl.x, q[1] * scl.y, q[2] * scl.z};   // scale, then translate
                    *p[k] = {w.x + pos.x, w.y + pos.y, w.z + pos.z};
                }
                sh.e = e;
                sh.c = c;
                sh.refl = refl_of(mat);
                sh.area = tri_area(sh.t.p1, sh.t.p2, sh.t.p3);
                if (!is_zero(e)) lights++;
                out.push_back(sh);
            }
        } else {
            PT_CHECK(false, PTGPU_ERR_INVALID, "%s: unknown object type '%s'", path, s.c_str());
        }
    }
    PT_CHECK(!out.empty(), PTGPU_ERR_INVALID, "%s: no shapes", path);
    Shape *buf = static_cast<Shape *>(std::malloc(out.size() * sizeof(Shape)));
    PT_CHECK(buf, PTGPU_ERR_OOM, "out of host memory");
    std::memcpy(buf, out.data(), out.size() * sizeof(Shape));
    *h_shapes = buf;
    *shapeCnt = (unsigned)out.size();
    *lightCnt = lights;
    return PTGPU_OK;
}

int ptgpu_write_ppm(const char *path, const unsigned int *h_pixels, int width, int height) {
    PT_CHECK(path && h_pixels && width > 0 && height > 0, PTGPU_ERR_INVALID, "bad argument");
    FILE *f = std::fopen(path, "w");
    PT_CHECK(f, PTGPU_ERR_INVALID, "cannot open %s for writing", path);
    std::fprintf(f, "P3\n%d %d\n%d\n", width, height, 255);
    for (int y = height - 1; y >= 0; --y) {
        const unsigned *row = h_pixels + (size_t)y * width;
        for (int x = 0; x < width; ++x)
            std::fprintf(f, "%u %u %u ", row[x] & 0xffu, (row[x] >> 8) & 0xffu, (row[x] >> 16) & 0xffu);
    }
    const bool ok = std::fclose(f) == 0;
    PT_CHECK(ok, PTGPU_ERR_INVALID, "write to %s failed", path);
    return PTGPU_OK;
}

}  // extern "C"
