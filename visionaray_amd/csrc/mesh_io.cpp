// mesh_io.cpp -- scene ingestion: Wavefront OBJ files (with their MTL
// material libraries) and PLY files read into the reference's model
// (src/common/model.cpp, obj_loader.cpp, obj_grammar.cpp, ply_loader.cpp).
//
// The reference parses a file in one thread with a Boost.Spirit grammar.
// Here an OBJ file is memory-mapped and cut at line ends into one piece per
// thread.  Pass 1 (threads): each piece's lines become vertices, texture
// coordinates, normals, raw face corners and the material directives, each
// face remembering how many vertices its piece had read before it.  Pass 2
// (serial, cheap): every piece learns how many vertices the pieces before it
// hold, and the directives are replayed in file order (mtllib reads a
// library, usemtl appends a material), which fixes the material id of every
// run of faces.  Pass 3 (threads): indices are resolved against the vertex
// count at the face, faces are split into triangle fans and zero-area
// triangles are dropped.  The pieces' triangles are then joined in file
// order, so the triangle list, its ids, normals and texture coordinates are
// those of the reference's single pass.
//
// The grammar is the reference's, line by line (obj_grammar.cpp:35-64):
//   `v` 3, 4 or 6 reals; `vt` 2 or 3 reals; `vn` 3 reals;
//   `f` 3 or more corners v[/[t][/[n]]] (no blanks inside a corner);
//   `usemtl` / `mtllib` / `#` take the rest of the line; anything else is
//   skipped.  Reals are parsed correctly rounded (std::from_chars) with an
//   optional leading '+' as qi::float_ allows.  A keyword is matched as a
//   prefix after leading blanks (space, tab), a line ends at "\r\n", "\n"
//   or "\r", and a final line without a line end matches no rule (every
//   rule ends in qi::eol), so it is skipped.
#include "mesh.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace vr {

void* huge_alloc(size_t bytes)
{
    constexpr size_t HUGE = (size_t)2 << 20;
    const size_t sz = (bytes + HUGE - 1) / HUGE * HUGE;
    void* p = std::aligned_alloc(HUGE, sz);
    if (!p) throw std::bad_alloc();
    madvise(p, sz, MADV_HUGEPAGE);
    return p;
}

void huge_free(void* p) { std::free(p); }

namespace {

// ---------------------------------------------------------------------------
// files
// ---------------------------------------------------------------------------

struct mapped_file {
    const char* data = "";
    size_t size = 0;
    void* map = nullptr;

    bool open(const std::string& path)
    {
        const int fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0) return false;
        struct stat sb;
        if (fstat(fd, &sb) != 0 || !S_ISREG(sb.st_mode)) {
            ::close(fd);
            return false;
        }
        size = (size_t)sb.st_size;
        if (size > 0) {
            map = mmap(nullptr, size, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
            if (map == MAP_FAILED) {
                map = nullptr;
                ::close(fd);
                return false;
            }
            data = static_cast<const char*>(map);
        }
        ::close(fd);
        return true;
    }
    ~mapped_file()
    {
        if (map) munmap(map, size);
    }
};

bool is_regular_file(const std::string& path)
{
    struct stat sb;
    return stat(path.c_str(), &sb) == 0 && S_ISREG(sb.st_mode);
}

// boost::filesystem::path(p).parent_path(): everything before the last '/'
std::string parent_path(const std::string& p)
{
    const size_t k = p.find_last_of('/');
    if (k == std::string::npos) return std::string();
    if (k == 0) return std::string("/");
    return p.substr(0, k);
}

std::string extension(const std::string& p)
{
    const size_t slash = p.find_last_of('/');
    const size_t dot = p.find_last_of('.');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return std::string();
    return p.substr(dot);
}

unsigned io_threads()
{
    unsigned n = std::max(1u, std::thread::hardware_concurrency());
    n = std::min(n, 16u);
    if (const char* s = std::getenv("VR_IO_THREADS")) {
        const long v = std::strtol(s, nullptr, 10);
        if (v >= 1 && v <= 256) n = (unsigned)v;
    }
    return n;
}

template <typename F>
void parallel_for(unsigned n, F&& f)
{
    if (n <= 1) {
        for (unsigned i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> ts;
    ts.reserve(n - 1);
    for (unsigned i = 1; i < n; ++i) ts.emplace_back([&f, i] { f(i); });
    f(0);
    for (auto& t : ts) t.join();
}

// ---------------------------------------------------------------------------
// lexical rules of the grammar
// ---------------------------------------------------------------------------

inline bool is_blank(char c) { return c == ' ' || c == '\t'; }

inline const char* skip_blanks(const char* p, const char* e)
{
    while (p < e && is_blank(*p)) ++p;
    return p;
}

// the first '\r' or '\n' at or after p (e if none)
inline const char* line_end(const char* p, const char* e)
{
    while (p < e && *p != '\n' && *p != '\r') ++p;
    return p;
}

// past qi::eol at le ("\r\n", "\r" or "\n")
inline const char* past_eol(const char* le, const char* e)
{
    if (*le == '\r' && le + 1 < e && le[1] == '\n') return le + 2;
    return le + 1;
}

inline bool starts_with(const char* p, const char* e, const char* kw, size_t n)
{
    return (size_t)(e - p) >= n && std::memcmp(p, kw, n) == 0;
}

// qi::float_: [+-] followed by a real (digits with an optional point and
// exponent, or nan / inf / infinity), correctly rounded to float.  Short
// decimals take Clinger's fast path (W. D. Clinger, "How to read floating
// point numbers accurately", PLDI 1990): an integer significand m <= 2^24 and
// a scale 10^k with |k| <= 10 are both exact floats, so one IEEE multiply or
// divide is the correctly rounded value.  Everything else (and nan / inf)
// goes to std::from_chars, which covers the same token.
constexpr float POW10F[11] = {1e0f, 1e1f, 1e2f, 1e3f, 1e4f, 1e5f, 1e6f, 1e7f, 1e8f, 1e9f, 1e10f};

inline bool is_digit(char c) { return (unsigned)(c - '0') < 10u; }

inline bool parse_real(const char*& p, const char* e, float& out)
{
    const char* s = p;
    bool neg = false;
    if (s < e && (*s == '+' || *s == '-')) {
        neg = *s == '-';
        ++s;
        if (s < e && (*s == '+' || *s == '-')) return false;
    }
    const char* q = s;
    uint64_t m = 0;
    int scale = 0;
    bool any = false, wide = false;
    for (; q < e && is_digit(*q); ++q) {
        any = true;
        if (m < 100000000000000000ull) m = m * 10 + (uint64_t)(*q - '0');
        else wide = true;
    }
    if (q < e && *q == '.') {
        for (++q; q < e && is_digit(*q); ++q) {
            any = true;
            if (m < 100000000000000000ull) {
                m = m * 10 + (uint64_t)(*q - '0');
                --scale;
            } else {
                wide = true;
            }
        }
    }
    if (any && !wide) {
        int ex = 0;
        const char* r = q;
        if (r < e && (*r == 'e' || *r == 'E')) {
            ++r;
            bool eneg = false;
            if (r < e && (*r == '+' || *r == '-')) eneg = *r++ == '-';
            if (r < e && is_digit(*r)) {
                for (; r < e && is_digit(*r); ++r)
                    if (ex < 100000) ex = ex * 10 + (*r - '0');
                q = r;
                ex = eneg ? -ex : ex;
            }
        }
        const int k = scale + ex;
        if (m <= (1u << 24) && k >= -10 && k <= 10) {
            const float v = k >= 0 ? (float)m * POW10F[k] : (float)m / POW10F[-k];
            out = neg ? -v : v;
            p = q;
            return true;
        }
    }
    float v = 0.0f;
    const auto r = std::from_chars(s, e, v, std::chars_format::general);
    if (r.ec != std::errc()) return false;
    out = neg ? -v : v;
    p = r.ptr;
    return true;
}

// qi::int_: [+-] digits, failing on 32-bit overflow
inline bool parse_int(const char*& p, const char* e, int32_t& out)
{
    const char* s = p;
    bool neg = false;
    if (s < e && (*s == '+' || *s == '-')) {
        neg = *s == '-';
        ++s;
    }
    if (s >= e || *s < '0' || *s > '9') return false;
    int64_t v = 0;
    while (s < e && *s >= '0' && *s <= '9') {
        v = v * 10 + (*s - '0');
        if (v > (int64_t)INT32_MAX + 1) return false;
        ++s;
    }
    if (!neg && v > INT32_MAX) return false;
    if (neg && v == (int64_t)INT32_MAX + 1) return false;  // INT32_MIN marks an absent index below
    out = (int32_t)(neg ? -v : v);
    p = s;
    return true;
}

// reals up to the line end: their count, or -1 if anything else is on the
// line or there are more than max of them
inline int parse_reals(const char* q, const char* le, float* out, int max)
{
    int k = 0;
    for (;;) {
        q = skip_blanks(q, le);
        if (q == le) return k;
        if (k == max) return -1;
        if (!parse_real(q, le, out[k])) return -1;
        ++k;
    }
}

inline std::string trimmed(const char* b, const char* e)
{
    auto space = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; };
    while (b < e && space(*b)) ++b;
    while (e > b && space(e[-1])) --e;
    return std::string(b, e);
}

// ---------------------------------------------------------------------------
// the model's arithmetic
// ---------------------------------------------------------------------------

inline f3 vertex(const float* pos, uint32_t i) { return f3{pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]}; }

// aabb::insert (math/detail/aabb.inl): component-wise min / max
inline void insert(f3& lo, f3& hi, f3 v)
{
    lo.x = v.x < lo.x ? v.x : lo.x;
    lo.y = v.y < lo.y ? v.y : lo.y;
    lo.z = v.z < lo.z ? v.z : lo.z;
    hi.x = hi.x < v.x ? v.x : hi.x;
    hi.y = hi.y < v.y ? v.y : hi.y;
    hi.z = hi.z < v.z ? v.z : hi.z;
}

vr_obj_material default_material()
{
    // sg::obj_material's defaults (src/common/sg/material.h:43-71)
    vr_obj_material m;
    for (int k = 0; k < 3; ++k) {
        m.ca[k] = 0.2f;
        m.cd[k] = 0.8f;
        m.cs[k] = 0.1f;
        m.ce[k] = 0.0f;
        m.ior[k] = 1.0f;
    }
    m.transmission = 0.0f;
    m.specular_exp = 32.0f;
    m.illum = 2;
    return m;
}

// ---------------------------------------------------------------------------
// MTL (parse_mtl, obj_loader.cpp:182-242; mtl defaults :163-175)
// ---------------------------------------------------------------------------

struct mtl {
    float ka[3] = {0.2f, 0.2f, 0.2f};
    float kd[3] = {0.8f, 0.8f, 0.8f};
    float ke[3] = {0.0f, 0.0f, 0.0f};
    float ks[3] = {0.1f, 0.1f, 0.1f};
    float tr = 0.0f;
    float d = 1.0f;
    float ns = 32.0f;
    float ni = 1.0f;
    std::string map_kd;
    int32_t illum = 2;
};

bool parse_mtl(const std::string& path, std::map<std::string, mtl>& lib)
{
    mapped_file f;
    if (!f.open(path)) return false;
    const char* p = f.data;
    const char* e = f.data + f.size;
    mtl* cur = nullptr;
    while (p < e) {
        const char* le = line_end(p, e);
        if (le == e) break;  // a final line without a line end (the reference would not terminate on it)
        const char* q = skip_blanks(p, le);
        float v[3];
        auto vec3 = [&](size_t n, float* dst) {
            if (parse_reals(q + n, le, v, 3) != 3) return false;
            dst[0] = v[0];
            dst[1] = v[1];
            dst[2] = v[2];
            return true;
        };
        auto real = [&](size_t n, float& dst) {
            if (parse_reals(q + n, le, v, 1) != 1) return false;
            dst = v[0];
            return true;
        };
        if (starts_with(q, le, "newmtl", 6)) {
            // std::map::insert keeps an existing entry: later lines edit it
            cur = &lib.insert({trimmed(q + 6, le), mtl()}).first->second;
        } else if (cur) {
            if (starts_with(q, le, "Ka", 2) && vec3(2, cur->ka)) {
            } else if (starts_with(q, le, "Kd", 2) && vec3(2, cur->kd)) {
            } else if (starts_with(q, le, "Ke", 2) && vec3(2, cur->ke)) {
            } else if (starts_with(q, le, "Ks", 2) && vec3(2, cur->ks)) {
            } else if (starts_with(q, le, "Tr", 2) && real(2, cur->tr)) {
            } else if (starts_with(q, le, "d", 1) && real(1, cur->d)) {
            } else if (starts_with(q, le, "Ns", 2) && real(2, cur->ns)) {
            } else if (starts_with(q, le, "Ni", 2) && real(2, cur->ni)) {
            } else if (starts_with(q, le, "map_Kd", 6)) {
                cur->map_kd.assign(skip_blanks(q + 6, le), le);
            } else if (starts_with(q, le, "illum", 5)) {
                const char* r = skip_blanks(q + 5, le);
                int32_t iv;
                if (parse_int(r, le, iv) && skip_blanks(r, le) == le) cur->illum = iv;
            }
        }
        p = past_eol(le, e);
    }
    return true;
}

// add_material (obj_loader.cpp:248-263)
vr_obj_material from_mtl(const mtl& m)
{
    vr_obj_material o;
    for (int k = 0; k < 3; ++k) {
        o.ca[k] = m.ka[k];
        o.cd[k] = m.kd[k];
        o.cs[k] = m.ks[k];
        o.ce[k] = m.ke[k];
        o.ior[k] = m.ni;
    }
    o.transmission = m.tr > 0.0f ? m.tr : 0.0f;
    o.transmission = 1.0f - m.d > 0.0f ? 1.0f - m.d : o.transmission;
    o.specular_exp = m.ns;
    o.illum = m.illum;
    return o;
}

// ---------------------------------------------------------------------------
// OBJ
// ---------------------------------------------------------------------------

constexpr int32_t NO_INDEX = INT32_MIN;

struct obj_piece {
    const char* b = nullptr;
    const char* e = nullptr;
    // pass 1
    pod_vector<float> v, vt, vn;
    pod_vector<int32_t> corner;  // 3 per corner: v, t, n (NO_INDEX when absent)
    struct face {
        uint64_t corner0;
        uint32_t n;
        uint32_t nv, nvt, nvn;  // the piece's counts before this face
        uint64_t offset;        // byte offset of the face's line in the file
    };
    std::vector<face> faces;
    struct directive {
        bool usemtl;
        std::string arg;
        uint64_t face;  // the piece's faces before it
    };
    std::vector<directive> dirs;
    // pass 2
    uint64_t vbase = 0, vtbase = 0, vnbase = 0;
    std::vector<std::pair<uint64_t, uint32_t>> runs;  // (first face, material id)
    // pass 3
    pod_vector<uint32_t> idx, geom;
    pod_vector<float> tex, shading;
    uint64_t degenerate = 0;
    f3 lo{FLT_MAX, FLT_MAX, FLT_MAX}, hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};  // over v1, v1 + e1, v1 + e2
    int err = 0;
    std::string msg;
};

void obj_pass1(obj_piece& pc, const char* file_begin)
{
    std::vector<int32_t> fc;
    const char* p = pc.b;
    const char* e = pc.e;
    while (p < e) {
        const char* le = line_end(p, e);
        if (le == e) break;  // unterminated final line: matches no rule
        const char* q = skip_blanks(p, le);
        if (q < le) {
            const char c = *q;
            if (c == 'v') {
                float r[7];
                int k = parse_reals(q + 1, le, r, 6);
                if (k == 3 || k == 4 || k == 6) {  // v x y z [w] | v x y z r g b
                    pc.v.insert(pc.v.end(), r, r + 3);
                } else if (starts_with(q, le, "vt", 2) && ((k = parse_reals(q + 2, le, r, 3)) == 2 || k == 3)) {
                    pc.vt.insert(pc.vt.end(), r, r + 2);
                } else if (starts_with(q, le, "vn", 2) && parse_reals(q + 2, le, r, 3) == 3) {
                    pc.vn.insert(pc.vn.end(), r, r + 3);
                }
            } else if (c == 'f') {
                // f idx idx idx {idx}; idx = int ['/'] [int] ['/'] [int], no blanks inside
                fc.clear();
                const char* s = q + 1;
                bool ok = true;
                for (;;) {
                    s = skip_blanks(s, le);
                    if (s == le) break;
                    int32_t vi, ti = NO_INDEX, ni = NO_INDEX;
                    if (!parse_int(s, le, vi)) {
                        ok = false;
                        break;
                    }
                    if (s < le && *s == '/') ++s;
                    parse_int(s, le, ti);
                    if (s < le && *s == '/') ++s;
                    parse_int(s, le, ni);
                    fc.push_back(vi);
                    fc.push_back(ti);
                    fc.push_back(ni);
                }
                if (ok && fc.size() >= 9) {
                    pc.faces.push_back({pc.corner.size() / 3, (uint32_t)(fc.size() / 3), (uint32_t)(pc.v.size() / 3),
                                        (uint32_t)(pc.vt.size() / 2), (uint32_t)(pc.vn.size() / 3),
                                        (uint64_t)(p - file_begin)});
                    pc.corner.insert(pc.corner.end(), fc.begin(), fc.end());
                }
            } else if (c == 'm' && starts_with(q, le, "mtllib", 6)) {
                pc.dirs.push_back({false, std::string(skip_blanks(q + 6, le), le), pc.faces.size()});
            } else if (c == 'u' && starts_with(q, le, "usemtl", 6)) {
                pc.dirs.push_back({true, trimmed(q + 6, le), pc.faces.size()});
            }
            // '#' comments and every other line: skipped
        }
        p = past_eol(le, e);
    }
}

// remap_index (obj_loader.cpp:40-44), checked against the count read so far
inline bool remap(int32_t i, uint64_t size, uint64_t& out)
{
    const int64_t r = i < 0 ? (int64_t)size + i : (int64_t)i - 1;
    if (r < 0 || (uint64_t)r >= size) return false;
    out = (uint64_t)r;
    return true;
}

size_t line_of(const char* file, uint64_t offset)
{
    return 1 + (size_t)std::count(file, file + offset, '\n');
}

// store_faces / store_triangle (obj_loader.cpp:51-132) for one piece
void obj_pass3(obj_piece& pc, const char* file, const std::string& path, const float* fpos, const float* ftex,
               const float* fnrm, uint64_t model_vbase)
{
    size_t run = 0;
    for (size_t fi = 0; fi < pc.faces.size(); ++fi) {
        const auto& f = pc.faces[fi];
        while (run + 1 < pc.runs.size() && pc.runs[run + 1].first <= fi) ++run;
        const uint32_t gid = pc.runs[run].second;
        const int32_t* c = pc.corner.data() + 3 * f.corner0;
        const uint64_t nv = pc.vbase + f.nv, nt = pc.vtbase + f.nvt, nn = pc.vnbase + f.nvn;
        auto bad = [&](const char* what) {
            pc.err = VR_ERR_INVALID_ARGUMENT;
            pc.msg = "obj: " + std::string(what) + " index out of range at " + path + ":" +
                     std::to_string(line_of(file, f.offset));
        };
        uint64_t i1, t1 = 0, n1 = 0;
        if (!remap(c[0], nv, i1)) return bad("vertex");
        for (uint32_t last = 2; last < f.n; ++last) {
            const int32_t* c2 = c + 3 * (last - 1);
            const int32_t* c3 = c + 3 * last;
            uint64_t i2, i3;
            if (!remap(c2[0], nv, i2) || !remap(c3[0], nv, i3)) return bad("vertex");
            const f3 v1 = vertex(fpos, (uint32_t)i1);
            const f3 e1 = sub3(vertex(fpos, (uint32_t)i2), v1);
            const f3 e2 = sub3(vertex(fpos, (uint32_t)i3), v1);
            const f3 cr = cross3(e1, e2);
            if (std::sqrt(dot3(cr, cr)) == 0.0f) {  // length(cross(e1, e2)) == 0
                ++pc.degenerate;  // rejected: zero area
                continue;
            }
            insert(pc.lo, pc.hi, v1);
            insert(pc.lo, pc.hi, add3(v1, e1));
            insert(pc.lo, pc.hi, add3(v1, e2));
            pc.idx.push_back((uint32_t)(model_vbase + i1));
            pc.idx.push_back((uint32_t)(model_vbase + i2));
            pc.idx.push_back((uint32_t)(model_vbase + i3));
            pc.geom.push_back(gid);
            if (c[1] != NO_INDEX && c2[1] != NO_INDEX && c3[1] != NO_INDEX) {
                uint64_t t2, t3;
                if (!remap(c[1], nt, t1) || !remap(c2[1], nt, t2) || !remap(c3[1], nt, t3))
                    return bad("texture coordinate");
                for (uint64_t t : {t1, t2, t3}) {
                    pc.tex.push_back(ftex[2 * t]);
                    pc.tex.push_back(ftex[2 * t + 1]);
                }
            }
            if (c[2] != NO_INDEX && c2[2] != NO_INDEX && c3[2] != NO_INDEX) {
                uint64_t n2, n3;
                if (!remap(c[2], nn, n1) || !remap(c2[2], nn, n2) || !remap(c3[2], nn, n3))
                    return bad("normal");
                for (uint64_t n : {n1, n2, n3}) {
                    pc.shading.push_back(fnrm[3 * n]);
                    pc.shading.push_back(fnrm[3 * n + 1]);
                    pc.shading.push_back(fnrm[3 * n + 2]);
                }
            }
        }
    }
}

struct obj_state {
    std::map<std::string, mtl> lib;
    std::vector<std::string> parsed_libs;
    uint64_t geom_id = 0;
    uint64_t first_prim = 0;  // triangles of the model before this load_obj call
    f3 lo{FLT_MAX, FLT_MAX, FLT_MAX}, hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};  // this call's triangles
};

inline void merge(f3& lo, f3& hi, f3 blo, f3 bhi)
{
    lo = f3{blo.x < lo.x ? blo.x : lo.x, blo.y < lo.y ? blo.y : lo.y, blo.z < lo.z ? blo.z : lo.z};
    hi = f3{hi.x < bhi.x ? bhi.x : hi.x, hi.y < bhi.y ? bhi.y : hi.y, hi.z < bhi.z ? bhi.z : hi.z};
}

template <typename T>
void append_parts(pod_vector<T>& dst, std::vector<obj_piece>& pcs, pod_vector<T> obj_piece::*part, unsigned nt)
{
    std::vector<size_t> off(pcs.size() + 1, dst.size());
    for (size_t k = 0; k < pcs.size(); ++k) off[k + 1] = off[k] + (pcs[k].*part).size();
    dst.resize(off.back());
    parallel_for(std::min<unsigned>(nt, (unsigned)pcs.size()), [&](unsigned t) {
        for (size_t k = t; k < pcs.size(); k += nt) {
            auto& src = pcs[k].*part;
            if (!src.empty()) std::memcpy(dst.data() + off[k], src.data(), src.size() * sizeof(T));
            pod_vector<T>().swap(src);
        }
    });
}

int load_obj_file(mesh_model& m, obj_state& S, const std::string& path)
{
    mapped_file f;
    VR_CHECK(f.open(path), VR_ERR_INVALID_ARGUMENT, "obj: cannot open " + path);
    const unsigned nt = io_threads();
    // pieces: about size / nt bytes each, every one ending just past a '\n'
    const unsigned want = f.size >= ((size_t)1 << 20) ? nt : 1u;
    std::vector<obj_piece> pcs(want);
    const char* e = f.data + f.size;
    const char* s = f.data;
    for (unsigned k = 0; k < want; ++k) {
        pcs[k].b = s;
        const char* t = k + 1 == want ? e : std::max(s, f.data + f.size / want * (k + 1));
        if (t < e) {
            const void* nl = std::memchr(t, '\n', (size_t)(e - t));
            t = nl ? static_cast<const char*>(nl) + 1 : e;
        }
        pcs[k].e = t;
        s = t;
    }
    parallel_for(want, [&](unsigned k) { obj_pass1(pcs[k], f.data); });

    // pass 2: vertex counts before each piece, directives in file order
    uint64_t nv = 0, nvt = 0, nvn = 0;
    for (auto& pc : pcs) {
        pc.vbase = nv;
        pc.vtbase = nvt;
        pc.vnbase = nvn;
        nv += pc.v.size() / 3;
        nvt += pc.vt.size() / 2;
        nvn += pc.vn.size() / 3;
    }
    VR_CHECK(m.pos.size() / 3 + nv <= 0xFFFFFFFFull, VR_ERR_UNSUPPORTED, "obj: more than 2^32 vertices");
    auto current_id = [&]() { return m.mats.empty() ? 0u : (uint32_t)(m.mats.size() - 1); };
    for (auto& pc : pcs) {
        pc.runs.push_back({0, current_id()});
        for (const auto& d : pc.dirs) {
            if (!d.usemtl) {  // mtllib (obj_loader.cpp:341-378): each name read once, next to the obj file
                if (std::find(S.parsed_libs.begin(), S.parsed_libs.end(), d.arg) != S.parsed_libs.end()) continue;
                const std::string dir = parent_path(path);
                const std::string lp = dir.empty() ? d.arg : dir + "/" + d.arg;
                if (!(is_regular_file(lp) && parse_mtl(lp, S.lib))) ++m.missing;
                S.parsed_libs.push_back(d.arg);
            } else {  // usemtl (obj_loader.cpp:379-469): a material appended per use of a known name
                const auto it = S.lib.find(d.arg);
                if (it != S.lib.end()) {
                    m.mats.push_back(from_mtl(it->second));
                    m.names.push_back(d.arg);
                } else {
                    ++m.missing;
                }
                S.geom_id = current_id();
                pc.runs.push_back({d.face, current_id()});
            }
        }
    }

    // vertex, texture coordinate and normal arrays of the whole file
    const uint64_t model_vbase = m.pos.size() / 3;
    append_parts(m.pos, pcs, &obj_piece::v, nt);
    pod_vector<float> ftex, fnrm;
    append_parts(ftex, pcs, &obj_piece::vt, nt);
    append_parts(fnrm, pcs, &obj_piece::vn, nt);
    const float* fpos = m.pos.data() + 3 * model_vbase;

    parallel_for(want, [&](unsigned k) {
        obj_pass3(pcs[k], f.data, path, fpos, ftex.data(), fnrm.data(), model_vbase);
    });
    for (auto& pc : pcs)
        if (pc.err) return fail(pc.err, pc.msg);

    const uint64_t first = m.geom.size();
    for (auto& pc : pcs) {
        m.degenerate += pc.degenerate;
        merge(S.lo, S.hi, pc.lo, pc.hi);
    }
    append_parts(m.idx, pcs, &obj_piece::idx, nt);
    append_parts(m.geom, pcs, &obj_piece::geom, nt);
    append_parts(m.tex, pcs, &obj_piece::tex, nt);
    append_parts(m.shading, pcs, &obj_piece::shading, nt);
    VR_CHECK(m.geom.size() <= 0x7FFFFFFFull, VR_ERR_UNSUPPORTED, "obj: more than 2^31 triangles");
    if (m.geom.size() > first) m.spans.push_back({first, m.geom.size() - first, true});

    // a material for every id used (obj_loader.cpp:493-497)
    while (m.mats.size() <= S.geom_id) {
        m.mats.push_back(default_material());
        m.names.emplace_back();
    }
    return VR_OK;
}

// after the files of one load_obj call (obj_loader.cpp:506-519): texture
// coordinates padded to three per triangle, bounds over v1, v1 + e1, v1 + e2
// of every triangle of the model (this call's were gathered by its pieces)
void finish_obj(mesh_model& m, const obj_state& S)
{
    const size_t nprims = m.geom.size();
    if (m.tex.size() < 6 * nprims) m.tex.resize(6 * nprims, 0.0f);
    merge(m.lo, m.hi, S.lo, S.hi);
    const float* pos = m.pos.data();
    for (size_t i = 0; i < S.first_prim; ++i) {
        const f3 v1 = vertex(pos, m.idx[3 * i]);
        const f3 e1 = sub3(vertex(pos, m.idx[3 * i + 1]), v1);
        const f3 e2 = sub3(vertex(pos, m.idx[3 * i + 2]), v1);
        insert(m.lo, m.hi, v1);
        insert(m.lo, m.hi, add3(v1, e1));
        insert(m.lo, m.hi, add3(v1, e2));
    }
}

// ---------------------------------------------------------------------------
// PLY (load_ply, ply_loader.cpp:64-202; the file format of tinyply)
// ---------------------------------------------------------------------------

enum ply_type { P_I8, P_U8, P_I16, P_U16, P_I32, P_U32, P_F32, P_F64, P_BAD };

ply_type ply_type_of(const std::string& s)
{
    if (s == "char" || s == "int8") return P_I8;
    if (s == "uchar" || s == "uint8") return P_U8;
    if (s == "short" || s == "int16") return P_I16;
    if (s == "ushort" || s == "uint16") return P_U16;
    if (s == "int" || s == "int32") return P_I32;
    if (s == "uint" || s == "uint32") return P_U32;
    if (s == "float" || s == "float32") return P_F32;
    if (s == "double" || s == "float64") return P_F64;
    return P_BAD;
}

inline size_t ply_size(ply_type t)
{
    static const size_t sz[] = {1, 1, 2, 2, 4, 4, 4, 8, 0};
    return sz[t];
}

struct ply_property {
    std::string name;
    ply_type type = P_BAD;
    bool list = false;
    ply_type count_type = P_BAD;
};

struct ply_element {
    std::string name;
    uint64_t count = 0;
    std::vector<ply_property> props;
    int find(const char* n) const
    {
        for (size_t k = 0; k < props.size(); ++k)
            if (props[k].name == n) return (int)k;
        return -1;
    }
};

// a cursor over the data section: binary (either byte order) or ASCII tokens
struct ply_reader {
    const char* p;
    const char* e;
    int format;  // 0 ascii, 1 little endian, 2 big endian
    bool ok = true;

    double value(ply_type t)
    {
        if (format == 0) {
            while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
            const char* s = p;
            while (p < e && !(*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
            if (s == p) {
                ok = false;
                return 0.0;
            }
            if (t == P_F32) {
                const char* q = s;
                float v;
                if (!parse_real(q, p, v) || q != p) ok = false;
                return v;
            }
            if (t == P_F64) {
                double v = 0.0;
                const char* q = s + (*s == '+' ? 1 : 0);
                const auto r = std::from_chars(q, p, v, std::chars_format::general);
                if (r.ec != std::errc() || r.ptr != p) ok = false;
                return v;
            }
            const char* q = s;
            int64_t v = 0;
            if (*q == '+') ++q;
            const auto r = std::from_chars(q, p, v);
            if (r.ec != std::errc() || r.ptr != p) ok = false;
            return (double)v;
        }
        const size_t n = ply_size(t);
        if ((size_t)(e - p) < n) {
            ok = false;
            return 0.0;
        }
        unsigned char b[8];
        std::memcpy(b, p, n);
        p += n;
        if (format == 2) std::reverse(b, b + n);
        switch (t) {
        case P_I8: return (double)(int8_t)b[0];
        case P_U8: return (double)b[0];
        case P_I16: { int16_t v; std::memcpy(&v, b, 2); return v; }
        case P_U16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
        case P_I32: { int32_t v; std::memcpy(&v, b, 4); return v; }
        case P_U32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
        case P_F32: { float v; std::memcpy(&v, b, 4); return v; }
        case P_F64: { double v; std::memcpy(&v, b, 8); return v; }
        default: ok = false; return 0.0;
        }
    }
    // a float property kept as its float value (a double one rounded once)
    float real(ply_type t) { return (float)value(t); }
};

int load_ply_file(mesh_model& m, const std::string& path)
{
    mapped_file f;
    VR_CHECK(f.open(path), VR_ERR_INVALID_ARGUMENT, "ply: cannot open " + path);
    const char* p = f.data;
    const char* e = f.data + f.size;
    // header: one keyword line each, up to end_header
    std::vector<ply_element> elems;
    int format = -1;
    bool magic = false, ended = false;
    while (p < e && !ended) {
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(e - p)));
        const char* le = nl ? nl : e;
        std::vector<std::string> tok;
        for (const char* s = p; s < le;) {
            while (s < le && (*s == ' ' || *s == '\t' || *s == '\r')) ++s;
            const char* t = s;
            while (s < le && !(*s == ' ' || *s == '\t' || *s == '\r')) ++s;
            if (s > t) tok.emplace_back(t, s);
        }
        p = nl ? nl + 1 : e;
        if (tok.empty()) continue;
        const std::string& k = tok[0];
        if (!magic) {
            VR_CHECK(k == "ply", VR_ERR_INVALID_ARGUMENT, "ply: " + path + " does not start with 'ply'");
            magic = true;
        } else if (k == "format") {
            VR_CHECK(tok.size() >= 2, VR_ERR_INVALID_ARGUMENT, "ply: bad format line in " + path);
            format = tok[1] == "ascii" ? 0 : tok[1] == "binary_little_endian" ? 1 : tok[1] == "binary_big_endian" ? 2 : -1;
            VR_CHECK(format >= 0, VR_ERR_INVALID_ARGUMENT, "ply: unknown format " + tok[1]);
        } else if (k == "element") {
            VR_CHECK(tok.size() >= 3, VR_ERR_INVALID_ARGUMENT, "ply: bad element line in " + path);
            ply_element el;
            el.name = tok[1];
            el.count = std::strtoull(tok[2].c_str(), nullptr, 10);
            elems.push_back(el);
        } else if (k == "property") {
            VR_CHECK(!elems.empty(), VR_ERR_INVALID_ARGUMENT, "ply: property before any element in " + path);
            ply_property pr;
            if (tok.size() >= 5 && tok[1] == "list") {
                pr.list = true;
                pr.count_type = ply_type_of(tok[2]);
                pr.type = ply_type_of(tok[3]);
                pr.name = tok[4];
                VR_CHECK(pr.count_type != P_BAD && pr.count_type != P_F32 && pr.count_type != P_F64,
                         VR_ERR_INVALID_ARGUMENT, "ply: bad list count type in " + path);
            } else {
                VR_CHECK(tok.size() >= 3, VR_ERR_INVALID_ARGUMENT, "ply: bad property line in " + path);
                pr.type = ply_type_of(tok[1]);
                pr.name = tok[2];
            }
            VR_CHECK(pr.type != P_BAD, VR_ERR_INVALID_ARGUMENT, "ply: unknown property type in " + path);
            elems.back().props.push_back(pr);
        } else if (k == "end_header") {
            ended = true;
        }
        // comment, obj_info: skipped
    }
    VR_CHECK(magic && ended && format >= 0, VR_ERR_INVALID_ARGUMENT, "ply: incomplete header in " + path);

    // the properties load_ply requests (ply_loader.cpp:84-102): vertex x y z
    // and face vertex_indices must exist; nx ny nz are optional, and u v are
    // read only when the normals are there (the same try block requests both)
    const ply_element* ve = nullptr;
    const ply_element* fe = nullptr;
    for (const auto& el : elems) {
        if (el.name == "vertex" && !ve) ve = &el;
        if (el.name == "face" && !fe) fe = &el;
    }
    VR_CHECK(ve && fe, VR_ERR_INVALID_ARGUMENT, "ply: " + path + " has no vertex or face element");
    const int px = ve->find("x"), py = ve->find("y"), pz = ve->find("z");
    const int fi = fe->find("vertex_indices");
    VR_CHECK(px >= 0 && py >= 0 && pz >= 0, VR_ERR_INVALID_ARGUMENT, "ply: vertex element lacks x, y or z");
    VR_CHECK(fi >= 0 && fe->props[fi].list, VR_ERR_INVALID_ARGUMENT, "ply: face element lacks vertex_indices");
    const int pnx = ve->find("nx"), pny = ve->find("ny"), pnz = ve->find("nz");
    const bool has_n = pnx >= 0 && pny >= 0 && pnz >= 0;
    const int pu = ve->find("u"), pv = ve->find("v");
    const bool has_uv = has_n && pu >= 0 && pv >= 0;
    for (int k : {px, py, pz})
        VR_CHECK(!ve->props[k].list && (ve->props[k].type == P_F32 || ve->props[k].type == P_F64),
                 VR_ERR_UNSUPPORTED, "ply: vertex positions must be float");
    VR_CHECK(fe->props[fi].type != P_F32 && fe->props[fi].type != P_F64, VR_ERR_INVALID_ARGUMENT,
             "ply: vertex_indices must be integers");
    VR_CHECK(m.pos.size() / 3 + ve->count <= 0xFFFFFFFFull, VR_ERR_UNSUPPORTED, "ply: more than 2^32 vertices");

    const uint64_t vbase = m.pos.size() / 3;
    pod_vector<float> nrm, uv;
    pod_vector<int64_t> fidx;
    ply_reader rd{p, e, format};
    const unsigned nt = io_threads();
    for (const auto& el : elems) {
        const bool is_v = &el == ve, is_f = &el == fe;
        if (is_v) {
            m.pos.resize(3 * (vbase + el.count));
            if (has_n) nrm.resize(3 * el.count);
            if (has_uv) uv.resize(2 * el.count);
        }
        if (is_f) fidx.resize(3 * el.count);
        // where each value of vertex r goes: slot 0-2 position, 3-5 normal, 6-7 u v
        auto store = [&](uint64_t r, int slot, float v) {
            if (slot < 3) m.pos[3 * (vbase + r) + slot] = v;
            else if (slot < 6) nrm[3 * r + slot - 3] = v;
            else uv[2 * r + slot - 6] = v;
        };
        std::vector<int> slot(el.props.size(), -1);
        if (is_v) {
            const int ks[8] = {px, py, pz, pnx, pny, pnz, pu, pv};
            for (int j = 0; j < 8; ++j)
                if (ks[j] >= 0 && (j < 3 || (j < 6 ? has_n : has_uv))) slot[ks[j]] = j;
        }
        // binary records of fixed size (every list but the face's index
        // list, read as 3 indices, makes the size variable) are decoded by
        // all threads
        size_t rec = 0;
        bool fixed = format != 0;
        std::vector<size_t> off(el.props.size());
        for (size_t k = 0; k < el.props.size(); ++k) {
            const auto& pr = el.props[k];
            off[k] = rec;
            if (!pr.list) rec += ply_size(pr.type);
            else if (is_f && (int)k == fi) rec += ply_size(pr.count_type) + 3 * ply_size(pr.type);
            else fixed = false;
        }
        if (fixed && rec > 0) {
            VR_CHECK(el.count <= (uint64_t)(e - rd.p) / rec, VR_ERR_INVALID_ARGUMENT, "ply: truncated data in " + path);
            const char* base = rd.p;
            std::vector<uint64_t> bad(nt, UINT64_MAX);
            const uint64_t per = (el.count + nt - 1) / nt;
            parallel_for(nt, [&](unsigned t) {
                ply_reader q{nullptr, nullptr, format};
                for (uint64_t r = t * per; r < std::min(el.count, (t + 1) * per); ++r) {
                    const char* at = base + r * rec;
                    for (size_t k = 0; k < el.props.size(); ++k) {
                        const auto& pr = el.props[k];
                        if (is_f && (int)k == fi) {
                            q.p = at + off[k];
                            q.e = at + rec;
                            if (q.value(pr.count_type) != 3.0) {
                                bad[t] = std::min(bad[t], r);
                                break;
                            }
                            for (int j = 0; j < 3; ++j) fidx[3 * r + j] = (int64_t)q.value(pr.type);
                        } else if (slot[k] >= 0) {
                            q.p = at + off[k];
                            q.e = at + rec;
                            store(r, slot[k], q.real(pr.type));
                        }
                    }
                    if (bad[t] != UINT64_MAX) break;
                }
            });
            const uint64_t b = *std::min_element(bad.begin(), bad.end());
            VR_CHECK(b == UINT64_MAX, VR_ERR_UNSUPPORTED,
                     "ply: face " + std::to_string(b) + " is not a triangle (load_ply reads 3 indices per face)");
            rd.p += el.count * rec;
            continue;
        }
        for (uint64_t r = 0; r < el.count && rd.ok; ++r) {
            for (size_t k = 0; k < el.props.size(); ++k) {
                const auto& pr = el.props[k];
                if (pr.list) {
                    const double cnt = rd.value(pr.count_type);
                    VR_CHECK(rd.ok && cnt >= 0.0, VR_ERR_INVALID_ARGUMENT, "ply: truncated data in " + path);
                    if (is_f && (int)k == fi) {
                        VR_CHECK(cnt == 3.0, VR_ERR_UNSUPPORTED,
                                 "ply: face " + std::to_string(r) + " is not a triangle (load_ply reads 3 indices per face)");
                        for (int j = 0; j < 3; ++j) fidx[3 * r + j] = (int64_t)rd.value(pr.type);
                    } else {
                        for (uint64_t j = 0; j < (uint64_t)cnt; ++j) rd.value(pr.type);
                    }
                } else if (slot[k] >= 0) {
                    store(r, slot[k], rd.real(pr.type));
                } else {
                    rd.value(pr.type);
                }
            }
        }
        VR_CHECK(rd.ok, VR_ERR_INVALID_ARGUMENT, "ply: truncated or malformed data in " + path);
    }

    // triangles in face order (ply_loader.cpp:120-176)
    const uint64_t nf = fidx.size() / 3;
    const uint64_t first = m.geom.size();
    VR_CHECK(first + nf <= 0x7FFFFFFFull, VR_ERR_UNSUPPORTED, "ply: more than 2^31 triangles");
    const uint32_t gid = (uint32_t)m.mats.size();
    const size_t i0 = m.idx.size(), s0 = m.shading.size(), t0 = m.tex.size();
    m.idx.resize(i0 + 3 * nf);
    m.geom.resize(first + nf);
    if (has_n) m.shading.resize(s0 + 9 * nf);
    m.tex.resize(t0 + 6 * nf);
    std::vector<uint64_t> bad(nt, UINT64_MAX);
    std::vector<f3> blo(nt, f3{FLT_MAX, FLT_MAX, FLT_MAX}), bhi(nt, f3{-FLT_MAX, -FLT_MAX, -FLT_MAX});
    const uint64_t per = (nf + nt - 1) / nt;
    parallel_for(nt, [&](unsigned t) {
        f3 lo{FLT_MAX, FLT_MAX, FLT_MAX}, hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint64_t i = t * per; i < std::min(nf, (t + 1) * per); ++i) {
            for (int j = 0; j < 3; ++j) {
                const int64_t ix = fidx[3 * i + j];
                if (ix < 0 || (uint64_t)ix >= ve->count) {
                    bad[t] = i;
                    return;
                }
                m.idx[i0 + 3 * i + j] = (uint32_t)(vbase + (uint64_t)ix);
                if (has_n)
                    for (int c = 0; c < 3; ++c) m.shading[s0 + 9 * i + 3 * j + c] = nrm[3 * ix + c];
                m.tex[t0 + 6 * i + 2 * j] = has_uv ? uv[2 * ix] : 0.0f;
                m.tex[t0 + 6 * i + 2 * j + 1] = has_uv ? uv[2 * ix + 1] : 0.0f;
                insert(lo, hi, vertex(m.pos.data(), (uint32_t)(vbase + (uint64_t)ix)));
            }
            m.geom[first + i] = gid;
        }
        blo[t] = lo;
        bhi[t] = hi;
    });
    const uint64_t b = *std::min_element(bad.begin(), bad.end());
    VR_CHECK(b == UINT64_MAX, VR_ERR_INVALID_ARGUMENT,
             "ply: vertex index out of range in face " + std::to_string(b) + " of " + path);
    for (unsigned t = 0; t < nt; ++t) merge(m.lo, m.hi, blo[t], bhi[t]);
    if (nf) m.spans.push_back({first, nf, false});
    m.mats.push_back(default_material());
    m.names.emplace_back();
    return VR_OK;
}

}  // namespace

// the model's geometric normal of triangle i (host side of vr_mesh_geometric_normals)
static f3 geometric_normal(const mesh_model& m, size_t i, bool normalized)
{
    const f3 v1 = vertex(m.pos.data(), m.idx[3 * i]);
    const f3 e1 = sub3(vertex(m.pos.data(), m.idx[3 * i + 1]), v1);
    const f3 e2 = sub3(vertex(m.pos.data(), m.idx[3 * i + 2]), v1);
    const f3 c = cross3(e1, e2);
    return normalized ? normalize3(c) : c;
}

}  // namespace vr

struct vr_mesh_t {
    vr::mesh_model m;
};

using namespace vr;

extern "C" {

int vr_mesh_load(const char* const* paths, size_t num_paths, vr_mesh* out)
{
    VR_TRY_BEGIN
    VR_CHECK(out && (num_paths == 0 || paths), VR_ERR_INVALID_ARGUMENT, "vr_mesh_load: null argument");
    *out = nullptr;
    VR_CHECK(num_paths > 0, VR_ERR_INVALID_ARGUMENT, "vr_mesh_load: no files");
    auto mesh = std::make_unique<vr_mesh_t>();
    mesh_model& m = mesh->m;
    // model::load (model.cpp:144-200): files of one format go to one loader
    // call (load_obj keeps its material library and material id across its
    // files); a list of mixed formats loads file by file
    auto type_of = [](const std::string& ext) {
        return ext == ".obj" || ext == ".OBJ" ? 1 : ext == ".ply" || ext == ".PLY" ? 2 : 0;
    };
    bool same = true;
    for (size_t k = 0; k < num_paths; ++k) {
        VR_CHECK(paths[k], VR_ERR_INVALID_ARGUMENT, "vr_mesh_load: null path");
        same = same && type_of(extension(paths[k])) == type_of(extension(paths[0]));
    }
    obj_state S;
    bool in_obj = false;
    for (size_t k = 0; k < num_paths; ++k) {
        const std::string path(paths[k]);
        const std::string ext = extension(path);
        const bool obj = type_of(ext) == 1;
        const bool ply = type_of(ext) == 2;
        VR_CHECK(obj || ply, VR_ERR_UNSUPPORTED, "vr_mesh_load: unsupported model format '" + ext + "' (" + path + ")");
        if (obj && !in_obj) {
            S = obj_state();
            S.first_prim = m.geom.size();
            in_obj = true;
        }
        const int rc = obj ? load_obj_file(m, S, path) : load_ply_file(m, path);
        if (rc) return rc;
        if (obj && !same) {
            finish_obj(m, S);
            in_obj = false;
        }
    }
    if (in_obj) finish_obj(m, S);
    *out = mesh.release();
    return VR_OK;
    VR_TRY_END
}

int vr_mesh_get_info(vr_mesh mesh, vr_mesh_info* info)
{
    VR_CHECK(mesh && info, VR_ERR_INVALID_ARGUMENT, "vr_mesh_get_info: null argument");
    const mesh_model& m = mesh->m;
    info->num_prims = m.geom.size();
    info->num_vertices = m.pos.size() / 3;
    info->num_shading_normals = m.shading.size() / 3;
    info->num_tex_coords = m.tex.size() / 2;
    info->num_materials = m.mats.size();
    info->num_degenerate = m.degenerate;
    info->num_missing = m.missing;
    info->bbox_lo[0] = m.lo.x;
    info->bbox_lo[1] = m.lo.y;
    info->bbox_lo[2] = m.lo.z;
    info->bbox_hi[0] = m.hi.x;
    info->bbox_hi[1] = m.hi.y;
    info->bbox_hi[2] = m.hi.z;
    return VR_OK;
}

int vr_mesh_triangles(vr_mesh mesh, vr_triangle* h_prims)
{
    VR_CHECK(mesh && (h_prims || mesh->m.geom.empty()), VR_ERR_INVALID_ARGUMENT, "vr_mesh_triangles: null argument");
    const mesh_model& m = mesh->m;
    for (size_t i = 0; i < m.geom.size(); ++i) {
        const f3 v1 = vertex(m.pos.data(), m.idx[3 * i]);
        const f3 e1 = sub3(vertex(m.pos.data(), m.idx[3 * i + 1]), v1);
        const f3 e2 = sub3(vertex(m.pos.data(), m.idx[3 * i + 2]), v1);
        vr_triangle& t = h_prims[i];
        t.geom_id = m.geom[i];
        t.prim_id = (uint32_t)i;
        t.v1[0] = v1.x, t.v1[1] = v1.y, t.v1[2] = v1.z;
        t.e1[0] = e1.x, t.e1[1] = e1.y, t.e1[2] = e1.z;
        t.e2[0] = e2.x, t.e2[1] = e2.y, t.e2[2] = e2.z;
    }
    return VR_OK;
}

int vr_mesh_geometric_normals(vr_mesh mesh, float* h_normals)
{
    VR_CHECK(mesh && (h_normals || mesh->m.geom.empty()), VR_ERR_INVALID_ARGUMENT,
             "vr_mesh_geometric_normals: null argument");
    const mesh_model& m = mesh->m;
    for (const mesh_span& s : m.spans)
        for (uint64_t i = s.first; i < s.first + s.count; ++i) {
            const f3 n = geometric_normal(m, i, s.normalized);
            h_normals[3 * i] = n.x;
            h_normals[3 * i + 1] = n.y;
            h_normals[3 * i + 2] = n.z;
        }
    return VR_OK;
}

int vr_mesh_shading_normals(vr_mesh mesh, float* h_normals)
{
    VR_CHECK(mesh && (h_normals || mesh->m.shading.empty()), VR_ERR_INVALID_ARGUMENT,
             "vr_mesh_shading_normals: null argument");
    if (!mesh->m.shading.empty())
        std::memcpy(h_normals, mesh->m.shading.data(), mesh->m.shading.size() * sizeof(float));
    return VR_OK;
}

int vr_mesh_tex_coords(vr_mesh mesh, float* h_tex_coords)
{
    VR_CHECK(mesh && (h_tex_coords || mesh->m.tex.empty()), VR_ERR_INVALID_ARGUMENT,
             "vr_mesh_tex_coords: null argument");
    if (!mesh->m.tex.empty()) std::memcpy(h_tex_coords, mesh->m.tex.data(), mesh->m.tex.size() * sizeof(float));
    return VR_OK;
}

int vr_mesh_materials(vr_mesh mesh, vr_obj_material* h_materials)
{
    VR_CHECK(mesh && (h_materials || mesh->m.mats.empty()), VR_ERR_INVALID_ARGUMENT,
             "vr_mesh_materials: null argument");
    if (!mesh->m.mats.empty())
        std::memcpy(h_materials, mesh->m.mats.data(), mesh->m.mats.size() * sizeof(vr_obj_material));
    return VR_OK;
}

int vr_mesh_material_name(vr_mesh mesh, size_t i, char* buf, size_t cap)
{
    VR_CHECK(mesh && buf && cap > 0, VR_ERR_INVALID_ARGUMENT, "vr_mesh_material_name: null argument");
    VR_CHECK(i < mesh->m.names.size(), VR_ERR_INVALID_ARGUMENT, "vr_mesh_material_name: index out of range");
    const std::string& s = mesh->m.names[i];
    const size_t n = std::min(cap - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
    return VR_OK;
}

int vr_mesh_pt_materials(vr_mesh mesh, int diffuse_fallback, vr_material* h_materials)
{
    VR_CHECK(mesh && (h_materials || mesh->m.mats.empty()), VR_ERR_INVALID_ARGUMENT,
             "vr_mesh_pt_materials: null argument");
    const auto& mats = mesh->m.mats;
    for (size_t i = 0; i < mats.size(); ++i) {
        const vr_obj_material& o = mats[i];
        vr_material& r = h_materials[i];
        const f3 ce{o.ce[0], o.ce[1], o.ce[2]};
        if (std::sqrt(dot3(ce, ce)) > 0.0f) {  // length(ce) > 0: emissive, ce, ls = 1
            r.type = 1;
            r.cd[0] = r.cd[1] = r.cd[2] = 0.0f;
            r.kd = 0.0f;
            r.ce[0] = o.ce[0], r.ce[1] = o.ce[1], r.ce[2] = o.ce[2];
            r.ls = 1.0f;
            continue;
        }
        VR_CHECK(o.illum == 1 || diffuse_fallback, VR_ERR_UNSUPPORTED,
                 "vr_mesh_pt_materials: material " + std::to_string(i) + " (illum " + std::to_string(o.illum) +
                     ") maps to a mirror, glass or plastic material, which the path tracer does not implement");
        r.type = 0;  // matte: cd, kd = 1
        r.cd[0] = o.cd[0], r.cd[1] = o.cd[1], r.cd[2] = o.cd[2];
        r.kd = 1.0f;
        r.ce[0] = r.ce[1] = r.ce[2] = 0.0f;
        r.ls = 0.0f;
    }
    return VR_OK;
}

int vr_mesh_upload(vr_mesh mesh, vr_triangle* d_prims, float* d_normals, void* stream)
{
    VR_TRY_BEGIN
    VR_CHECK(mesh && (d_prims || mesh->m.geom.empty()), VR_ERR_INVALID_ARGUMENT, "vr_mesh_upload: null argument");
    int rc = ensure_device();
    if (rc) return rc;
    return assemble_triangles(mesh->m, d_prims, d_normals, (hipStream_t)stream);
    VR_TRY_END
}

int vr_mesh_destroy(vr_mesh mesh)
{
    if (!mesh) return VR_OK;
    mesh_model& m = mesh->m;
    if (m.d_pos || m.d_idx || m.d_geom) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(m.device);
        (void)hipFree(m.d_pos);
        (void)hipFree(m.d_idx);
        (void)hipFree(m.d_geom);
        (void)hipSetDevice(cur);
    }
    delete mesh;
    return VR_OK;
}

}  // extern "C"
