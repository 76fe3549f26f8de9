// obj_loader.cpp — OBJ ingestion (include/bvh_obj.h).
//
// Same result as the reference's loadOBJ (src/loader.cpp:10-74): the line
// rules, std::istream float extraction, std::stoi face indices, negative
// (relative) indices, fan triangulation and the skip of triangles naming a
// vertex not read yet.  Faster: the file is read once into memory, split at
// line ends into one chunk per thread, and parsed in three parallel passes
// (vertices + raw face indices; face resolution against each chunk's vertex
// base; the SoA write at each chunk's triangle base), with prefix sums of
// the per-chunk counts between the passes.
#include "bvh_obj.h"

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <locale.h>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

locale_t c_locale() {
    static locale_t loc = newlocale(LC_ALL_MASK, "C", (locale_t)0);
    return loc;
}

inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }
inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

enum Extract { kOk, kFailed, kNone };

// One `stream >> float` on the characters [p, end) of a line, as libstdc++'s
// num_get does it in the "C" locale: skip white space; accumulate an
// optional sign, leading zeros (one kept), digits, at most one decimal point
// (not after an exponent), and an exponent 'e'/'E' (only after a mantissa
// digit) with an optional sign; convert the accumulated text with strtof.
// Text that strtof does not consume entirely gives 0 and failure; overflow
// gives +-FLT_MAX and failure.  kNone: the line ended before any character
// (the stream's sentry fails and the value is left alone).
Extract extract_float(const char*& p, const char* end, float& v, std::string& acc) {
    while (p < end && is_ws(*p)) ++p;
    if (p >= end) return kNone;
    acc.clear();
    bool mantissa = false;
    if (*p == '+' || *p == '-') acc += *p++;
    while (p < end && *p == '0') {
        if (!mantissa) {
            acc += '0';
            mantissa = true;
        }
        ++p;
    }
    bool dec = false, sci = false;
    while (p < end) {
        const char c = *p;
        if (c == '.' && !dec && !sci) {
            acc += '.';
            dec = true;
        } else if (is_digit(c)) {
            acc += c;
            mantissa = true;
        } else if ((c == 'e' || c == 'E') && !sci && mantissa) {
            acc += 'e';
            sci = true;
            ++p;
            if (p >= end) break;
            if (*p == '+' || *p == '-') acc += *p;
            else continue;  // the character after 'e' is examined again
        } else {
            break;
        }
        ++p;
    }
    char* stop = nullptr;
    const char* s = acc.c_str();
    const float x = strtof_l(s, &stop, c_locale());
    if (stop == s || *stop != '\0') {
        v = 0.0f;
        return kFailed;
    }
    if (x == std::numeric_limits<float>::infinity()) {
        v = FLT_MAX;
        return kFailed;
    }
    if (x == -std::numeric_limits<float>::infinity()) {
        v = -FLT_MAX;
        return kFailed;
    }
    v = x;
    return kOk;
}

// std::stoi of the token's text before its first '/': optional sign and
// decimal digits (trailing characters ignored).  False when there is no
// digit or the value does not fit an int (stoi throws).
bool face_index(const char* t, const char* tend, int& out) {
    const char* slash = static_cast<const char*>(std::memchr(t, '/', (size_t)(tend - t)));
    const char* e = slash ? slash : tend;
    const char* q = t;
    bool neg = false;
    if (q < e && (*q == '+' || *q == '-')) neg = *q++ == '-';
    if (q >= e || !is_digit(*q)) return false;
    long long v = 0;
    for (; q < e && is_digit(*q); ++q) {
        v = v * 10 + (*q - '0');
        if (v > (long long)INT_MAX + 1) return false;
    }
    if (neg) v = -v;
    if (v < INT_MIN || v > INT_MAX) return false;
    out = (int)v;
    return true;
}

struct Chunk {
    const char* b = nullptr;
    const char* e = nullptr;
    std::vector<float> verts;        // 3 per vertex
    std::vector<int> faces;          // per face: count, vertices read before it in this chunk, raw indices
    uint64_t line0 = 0;              // lines before this chunk
    uint64_t err_line = 0;           // 1-based line of the first bad face token (0 = none)
    uint64_t vbase = 0;              // vertices before this chunk
    std::vector<uint32_t> tris;      // 3 resolved vertex indices per triangle
    uint64_t tbase = 0;              // triangles before this chunk
};

void parse_chunk(Chunk& c) {
    std::string acc;
    acc.reserve(64);
    std::vector<int> idx;
    uint64_t line = c.line0;
    const char* p = c.b;
    while (p < c.e) {
        const char* le = static_cast<const char*>(std::memchr(p, '\n', (size_t)(c.e - p)));
        if (!le) le = c.e;
        ++line;
        const char* q = p;
        const char* lend = le;
        p = le < c.e ? le + 1 : le;
        if (q == lend || *q == '#') continue;
        while (q < lend && is_ws(*q)) ++q;
        const char* t0 = q;
        while (q < lend && !is_ws(*q)) ++q;
        const size_t plen = (size_t)(q - t0);
        if (plen == 1 && *t0 == 'v') {
            float xyz[3] = {0.0f, 0.0f, 0.0f};
            for (int k = 0; k < 3; ++k) {
                const Extract r = extract_float(q, lend, xyz[k], acc);
                if (r != kOk) break;  // a failed stream reads nothing more
            }
            c.verts.insert(c.verts.end(), xyz, xyz + 3);
        } else if (plen == 1 && *t0 == 'f') {
            idx.clear();
            while (true) {
                while (q < lend && is_ws(*q)) ++q;
                if (q >= lend) break;
                const char* ts = q;
                while (q < lend && !is_ws(*q)) ++q;
                int v = 0;
                if (!face_index(ts, q, v)) {
                    if (!c.err_line) c.err_line = line;
                    return;
                }
                idx.push_back(v);
            }
            c.faces.push_back((int)idx.size());
            c.faces.push_back((int)(c.verts.size() / 3));
            c.faces.insert(c.faces.end(), idx.begin(), idx.end());
        }
    }
}

void resolve_chunk(Chunk& c, uint64_t total_vertices_seen_limit) {
    (void)total_vertices_seen_limit;
    size_t k = 0;
    std::vector<long long> fi;
    while (k < c.faces.size()) {
        const int cnt = c.faces[k];
        const long long cur = (long long)c.vbase + c.faces[k + 1];  // vertices read before this face
        fi.resize((size_t)cnt);
        for (int j = 0; j < cnt; ++j) {
            const long long raw = c.faces[k + 2 + j];
            fi[(size_t)j] = raw < 0 ? cur + raw : raw - 1;
        }
        k += 2 + (size_t)cnt;
        for (int j = 1; j + 1 < cnt; ++j) {
            const long long a = fi[0], b = fi[(size_t)j], d = fi[(size_t)j + 1];
            if (a < 0 || a >= cur || b < 0 || b >= cur || d < 0 || d >= cur) continue;
            c.tris.push_back((uint32_t)a);
            c.tris.push_back((uint32_t)b);
            c.tris.push_back((uint32_t)d);
        }
    }
}

template <class F>
void parallel_for(int n, F f) {
    std::vector<std::thread> th;
    for (int i = 1; i < n; ++i) th.emplace_back(f, i);
    f(0);
    for (auto& t : th) t.join();
}

}  // namespace

struct bvh_obj_rec {
    std::vector<Chunk> chunks;
    std::vector<float> vertices;  // 3 per vertex, file order
    uint64_t nverts = 0;
    uint32_t ntri = 0;
};

extern "C" {

const char* bvh_obj_last_error(void) { return g_err.c_str(); }

int bvh_obj_parse(const char* path, int32_t threads, bvh_obj_t* out) {
    if (!path || !out) return BVH_E_ARG;
    *out = nullptr;
    g_err.clear();
    try {
        FILE* f = std::fopen(path, "rb");
        if (!f) {
            g_err = std::string("cannot open OBJ file: ") + path;
            return BVH_E_ARG;
        }
        std::string data;
        if (std::fseek(f, 0, SEEK_END) == 0) {
            const long sz = std::ftell(f);
            if (sz > 0) data.resize((size_t)sz);
            std::fseek(f, 0, SEEK_SET);
        }
        const size_t got = data.empty() ? 0 : std::fread(&data[0], 1, data.size(), f);
        std::fclose(f);
        data.resize(got);
        int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
        nt = std::max(1, std::min(nt, 64));
        if (data.size() < ((size_t)1 << 20)) nt = 1;
        auto obj = new bvh_obj_rec();
        std::unique_ptr<bvh_obj_rec> guard(obj);
        obj->chunks.resize((size_t)nt);
        // Chunk boundaries just after a line end.
        const char* base = data.data();
        const char* end = base + data.size();
        const char* cur = base;
        for (int i = 0; i < nt; ++i) {
            Chunk& c = obj->chunks[(size_t)i];
            c.b = cur;
            const char* want = i == nt - 1 ? end : base + (data.size() * (size_t)(i + 1)) / (size_t)nt;
            if (want < cur) want = cur;
            if (want < end) {
                const char* nl = static_cast<const char*>(std::memchr(want, '\n', (size_t)(end - want)));
                want = nl ? nl + 1 : end;
            }
            c.e = want;
            cur = want;
        }
        // Lines before each chunk (for error messages).
        parallel_for(nt, [&](int i) {
            Chunk& c = obj->chunks[(size_t)i];
            uint64_t l = 0;
            for (const char* p = c.b; p < c.e; ++p) l += *p == '\n';
            c.line0 = l;  // lines in this chunk, turned into a prefix below
        });
        uint64_t lines = 0;
        for (Chunk& c : obj->chunks) {
            const uint64_t l = c.line0;
            c.line0 = lines;
            lines += l;
        }
        parallel_for(nt, [&](int i) { parse_chunk(obj->chunks[(size_t)i]); });
        for (Chunk& c : obj->chunks)
            if (c.err_line) {
                g_err = std::string("OBJ face token without a vertex index at line ") + std::to_string(c.err_line);
                return BVH_E_ARG;
            }
        uint64_t vb = 0;
        for (Chunk& c : obj->chunks) {
            c.vbase = vb;
            vb += c.verts.size() / 3;
        }
        obj->nverts = vb;
        parallel_for(nt, [&](int i) { resolve_chunk(obj->chunks[(size_t)i], vb); });
        uint64_t tb = 0;
        for (Chunk& c : obj->chunks) {
            c.tbase = tb;
            tb += c.tris.size() / 3;
            std::vector<int>().swap(c.faces);
        }
        if (tb > (1ull << 30)) {
            g_err = "OBJ file has more than 2^30 triangles";
            return BVH_E_UNSUPPORTED;
        }
        obj->ntri = (uint32_t)tb;
        obj->vertices.resize((size_t)vb * 3);
        parallel_for(nt, [&](int i) {
            Chunk& c = obj->chunks[(size_t)i];
            std::copy(c.verts.begin(), c.verts.end(), obj->vertices.begin() + (ptrdiff_t)(c.vbase * 3));
            std::vector<float>().swap(c.verts);
        });
        *out = guard.release();
        return BVH_OK;
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory while parsing the OBJ file";
        return BVH_E_NOMEM;
    }
}

int bvh_obj_info(bvh_obj_t obj, uint32_t* triangles, uint64_t* vertices) {
    if (!obj) return BVH_E_ARG;
    if (triangles) *triangles = obj->ntri;
    if (vertices) *vertices = obj->nverts;
    return BVH_OK;
}

int bvh_obj_fill(bvh_obj_t obj, float* const* soa) {
    if (!obj || !soa) return BVH_E_ARG;
    for (int k = 0; k < 9; ++k)
        if (!soa[k] && obj->ntri) return BVH_E_ARG;
    const int nt = (int)obj->chunks.size();
    const float* V = obj->vertices.data();
    parallel_for(nt, [&](int i) {
        const Chunk& c = obj->chunks[(size_t)i];
        const size_t m = c.tris.size() / 3;
        for (size_t t = 0; t < m; ++t) {
            const size_t o = (size_t)c.tbase + t;
            for (int v = 0; v < 3; ++v) {
                const float* p = V + 3 * (size_t)c.tris[3 * t + (size_t)v];
                soa[3 * v + 0][o] = p[0];
                soa[3 * v + 1][o] = p[1];
                soa[3 * v + 2][o] = p[2];
            }
        }
    });
    return BVH_OK;
}

int bvh_obj_free(bvh_obj_t obj) {
    if (!obj) return BVH_E_ARG;
    delete obj;
    return BVH_OK;
}

int bvh_mesh_upload_obj(bvh_ctx_t ctx, const char* path, int32_t threads, bvh_mesh_t* out) {
    if (!ctx || !path || !out) return BVH_E_ARG;
    *out = nullptr;
    bvh_obj_t obj = nullptr;
    int st = bvh_obj_parse(path, threads, &obj);
    if (st != BVH_OK) return st;
    const uint32_t n = obj->ntri;
    if (n == 0) {
        bvh_obj_free(obj);
        g_err = "OBJ file has no triangles";
        return BVH_E_ARG;
    }
    float* staging = nullptr;
    if (bvh_host_alloc_pinned((size_t)n * 9 * sizeof(float), (void**)&staging) != BVH_OK) {
        bvh_obj_free(obj);
        g_err = "pinned staging allocation failed";
        return BVH_E_NOMEM;
    }
    float* soa[9];
    for (int k = 0; k < 9; ++k) soa[k] = staging + (size_t)k * n;
    st = bvh_obj_fill(obj, soa);
    bvh_obj_free(obj);
    if (st == BVH_OK) st = bvh_mesh_upload(ctx, soa, n, out);  // DMA from pinned memory
    (void)bvh_host_free_pinned(staging);
    return st;
}

}  // extern "C"
