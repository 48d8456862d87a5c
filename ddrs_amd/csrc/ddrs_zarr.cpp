// Zarr v3 reader for the binsparse-COO network stores (reference
// src/data/store/zarr.rs:47-300, which reads them with the zarrs crate): the
// arrays of one store group (order, indices_0/1, length_m, slope, optional
// parent_order / parent_offset) and a group's attributes (gage_idx,
// gage_catchment). Host code; the supported subset is what the stores use:
// regular chunk grids in C order, the "bytes" codec (either endian) optionally
// followed by "zstd", default chunk keys ("c/i/j" or "c.i.j") or v2 keys
// ("i.j"), missing chunks read as the fill value. zstd frames are decoded by the
// system's libzstd.so.1 through its documented C API, resolved at run time.
#include <dlfcn.h>

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ddrs_hip.h"

extern "C" void ddrs_internal_set_error(const char* msg);

namespace {

ddrs_status zfail(ddrs_status s, const std::string& m) {
    ddrs_internal_set_error(m.c_str());
    return s;
}

// --- a small JSON value and parser ------------------------------------------
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0;
    std::string str;  // STR, and the number's text for NUM
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;
    const JVal* get(const std::string& k) const {
        if (kind != OBJ) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

struct JParser {
    const char* p;
    const char* end;
    std::string err;
    void ws() {
        while (p < end && std::isspace(static_cast<unsigned char>(*p))) ++p;
    }
    bool lit(const char* s) {
        const size_t n = std::strlen(s);
        if (static_cast<size_t>(end - p) >= n && std::strncmp(p, s, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    bool str(std::string& out) {
        if (p >= end || *p != '"') return false;
        ++p;
        while (p < end && *p != '"') {
            if (*p == '\\') {
                if (++p >= end) return false;
                const char c = *p++;
                switch (c) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (end - p < 4) return false;
                        const unsigned cp = static_cast<unsigned>(std::strtoul(std::string(p, 4).c_str(), nullptr, 16));
                        p += 4;
                        if (cp < 0x80) out += static_cast<char>(cp);
                        else if (cp < 0x800) {
                            out += static_cast<char>(0xC0 | (cp >> 6));
                            out += static_cast<char>(0x80 | (cp & 0x3F));
                        } else {
                            out += static_cast<char>(0xE0 | (cp >> 12));
                            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
                            out += static_cast<char>(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: out += c;
                }
            } else {
                out += *p++;
            }
        }
        if (p >= end) return false;
        ++p;
        return true;
    }
    bool value(JVal& v, int depth) {
        if (depth > 64) return false;
        ws();
        if (p >= end) return false;
        if (*p == '{') {
            ++p;
            v.kind = JVal::OBJ;
            ws();
            if (p < end && *p == '}') { ++p; return true; }
            for (;;) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p >= end || *p != ':') return false;
                ++p;
                if (!value(v.obj[k], depth + 1)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == '}') { ++p; return true; }
                return false;
            }
        }
        if (*p == '[') {
            ++p;
            v.kind = JVal::ARR;
            ws();
            if (p < end && *p == ']') { ++p; return true; }
            for (;;) {
                v.arr.emplace_back();
                if (!value(v.arr.back(), depth + 1)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == ']') { ++p; return true; }
                return false;
            }
        }
        if (*p == '"') { v.kind = JVal::STR; return str(v.str); }
        if (lit("true")) { v.kind = JVal::BOOL; v.b = true; return true; }
        if (lit("false")) { v.kind = JVal::BOOL; return true; }
        if (lit("null")) { v.kind = JVal::NUL; return true; }
        const char* q = p;
        while (q < end && (std::isdigit(static_cast<unsigned char>(*q)) || *q == '-' || *q == '+' ||
                           *q == '.' || *q == 'e' || *q == 'E'))
            ++q;
        if (q == p) return false;
        v.kind = JVal::NUM;
        v.str.assign(p, q);
        v.num = std::strtod(v.str.c_str(), nullptr);
        p = q;
        return true;
    }
};

bool read_file(const std::string& path, std::vector<char>* out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out->resize(n > 0 ? static_cast<size_t>(n) : 0);
    const size_t got = n > 0 ? std::fread(out->data(), 1, out->size(), f) : 0;
    std::fclose(f);
    return got == out->size();
}

ddrs_status parse_json_file(const std::string& path, JVal* v) {
    std::vector<char> buf;
    if (!read_file(path, &buf)) return zfail(DDRS_ERR_INVALID, "zarr: cannot read " + path);
    JParser ps{buf.data(), buf.data() + buf.size(), {}};
    if (!ps.value(*v, 0)) return zfail(DDRS_ERR_INVALID, "zarr: malformed JSON in " + path);
    return DDRS_OK;
}

// --- zstd through the system library ----------------------------------------
struct Zstd {
    using decompress_t = size_t (*)(void*, size_t, const void*, size_t);
    using is_error_t = unsigned (*)(size_t);
    using err_name_t = const char* (*)(size_t);
    decompress_t decompress = nullptr;
    is_error_t is_error = nullptr;
    err_name_t err_name = nullptr;
    bool ok = false;
    Zstd() {
        void* h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) h = dlopen("libzstd.so", RTLD_NOW | RTLD_LOCAL);
        if (!h) return;
        decompress = reinterpret_cast<decompress_t>(dlsym(h, "ZSTD_decompress"));
        is_error = reinterpret_cast<is_error_t>(dlsym(h, "ZSTD_isError"));
        err_name = reinterpret_cast<err_name_t>(dlsym(h, "ZSTD_getErrorName"));
        ok = decompress && is_error && err_name;
    }
};

const Zstd& zstd() {
    static Zstd z;
    return z;
}

// --- arrays -------------------------------------------------------------------
struct DType {
    const char* name;
    int code;
    int size;
};
constexpr DType kTypes[] = {
    {"int8", DDRS_ZARR_I8, 1},     {"uint8", DDRS_ZARR_U8, 1},   {"int16", DDRS_ZARR_I16, 2},
    {"uint16", DDRS_ZARR_U16, 2},  {"int32", DDRS_ZARR_I32, 4},  {"uint32", DDRS_ZARR_U32, 4},
    {"int64", DDRS_ZARR_I64, 8},   {"uint64", DDRS_ZARR_U64, 8}, {"float32", DDRS_ZARR_F32, 4},
    {"float64", DDRS_ZARR_F64, 8},
};

struct ArrayMeta {
    std::vector<int64_t> shape, chunk;
    const DType* dt = nullptr;
    bool big_endian = false, zstd = false;
    std::string key_kind = "default";
    char sep = '/';
    std::vector<unsigned char> fill;  // one element, little endian
    int64_t count() const {
        int64_t n = 1;
        for (int64_t s : shape) n *= s;
        return n;
    }
};

std::string join(const std::string& a, const std::string& b) {
    if (a.empty()) return b;
    if (b.empty()) return a;
    return a.back() == '/' ? a + b : a + "/" + b;
}

ddrs_status fill_value(const JVal* f, const DType* dt, std::vector<unsigned char>* out) {
    out->assign(dt->size, 0);
    if (!f || f->kind == JVal::NUL) return DDRS_OK;
    double d = 0;
    if (f->kind == JVal::NUM) d = f->num;
    else if (f->kind == JVal::BOOL) d = f->b ? 1 : 0;
    else if (f->kind == JVal::STR && f->str == "NaN") d = NAN;
    else if (f->kind == JVal::STR && f->str == "Infinity") d = INFINITY;
    else if (f->kind == JVal::STR && f->str == "-Infinity") d = -INFINITY;
    else return zfail(DDRS_ERR_INVALID, "zarr: unsupported fill_value");
    switch (dt->code) {
        case DDRS_ZARR_F32: { const float v = static_cast<float>(d); std::memcpy(out->data(), &v, 4); break; }
        case DDRS_ZARR_F64: std::memcpy(out->data(), &d, 8); break;
        case DDRS_ZARR_U64: {
            const uint64_t v = f->kind == JVal::NUM ? std::strtoull(f->str.c_str(), nullptr, 10) : 0;
            std::memcpy(out->data(), &v, 8);
            break;
        }
        default: {
            const int64_t v = f->kind == JVal::NUM ? std::strtoll(f->str.c_str(), nullptr, 10) : static_cast<int64_t>(d);
            std::memcpy(out->data(), &v, dt->size);  // little-endian host
        }
    }
    return DDRS_OK;
}

ddrs_status read_meta(const std::string& dir, ArrayMeta* m) {
    JVal v;
    ddrs_status st = parse_json_file(join(dir, "zarr.json"), &v);
    if (st != DDRS_OK) return st;
    const JVal* zf = v.get("zarr_format");
    const JVal* nt = v.get("node_type");
    if (!zf || zf->num != 3 || !nt || nt->str != "array")
        return zfail(DDRS_ERR_INVALID, "zarr: " + dir + " is not a zarr v3 array");
    const JVal* shape = v.get("shape");
    const JVal* dtype = v.get("data_type");
    if (!shape || shape->kind != JVal::ARR || !dtype || dtype->kind != JVal::STR)
        return zfail(DDRS_ERR_INVALID, "zarr: " + dir + ": missing shape or data_type");
    for (const JVal& s : shape->arr) {
        if (s.kind != JVal::NUM || s.num < 0) return zfail(DDRS_ERR_INVALID, "zarr: bad shape");
        m->shape.push_back(static_cast<int64_t>(s.num));
    }
    for (const DType& t : kTypes)
        if (dtype->str == t.name) m->dt = &t;
    if (!m->dt) return zfail(DDRS_ERR_INVALID, "zarr: unsupported data_type " + dtype->str);
    const JVal* grid = v.get("chunk_grid");
    const JVal* gname = grid ? grid->get("name") : nullptr;
    const JVal* gconf = grid ? grid->get("configuration") : nullptr;
    const JVal* cs = gconf ? gconf->get("chunk_shape") : nullptr;
    if (!gname || gname->str != "regular" || !cs || cs->kind != JVal::ARR ||
        cs->arr.size() != m->shape.size())
        return zfail(DDRS_ERR_INVALID, "zarr: " + dir + ": only regular chunk grids are read");
    for (const JVal& s : cs->arr) {
        if (s.kind != JVal::NUM || s.num < 1) return zfail(DDRS_ERR_INVALID, "zarr: bad chunk shape");
        m->chunk.push_back(static_cast<int64_t>(s.num));
    }
    const JVal* codecs = v.get("codecs");
    if (!codecs || codecs->kind != JVal::ARR || codecs->arr.empty())
        return zfail(DDRS_ERR_INVALID, "zarr: " + dir + ": missing codecs");
    for (size_t i = 0; i < codecs->arr.size(); ++i) {
        const JVal& c = codecs->arr[i];
        const JVal* name = c.get("name");
        const std::string nm = name ? name->str : "";
        if (i == 0 && nm == "bytes") {
            const JVal* conf = c.get("configuration");
            const JVal* en = conf ? conf->get("endian") : nullptr;
            m->big_endian = en && en->str == "big";
        } else if (i == 1 && nm == "zstd") {
            m->zstd = true;
        } else {
            return zfail(DDRS_ERR_INVALID, "zarr: " + dir + ": unsupported codec chain (" + nm +
                                               "); this reader takes bytes [+ zstd]");
        }
    }
    if (const JVal* ke = v.get("chunk_key_encoding")) {
        const JVal* kn = ke->get("name");
        const JVal* conf = ke->get("configuration");
        const JVal* sep = conf ? conf->get("separator") : nullptr;
        m->key_kind = kn ? kn->str : "default";
        if (m->key_kind != "default" && m->key_kind != "v2")
            return zfail(DDRS_ERR_INVALID, "zarr: unsupported chunk_key_encoding " + m->key_kind);
        m->sep = m->key_kind == "v2" ? '.' : '/';
        if (sep && !sep->str.empty()) m->sep = sep->str[0];
    }
    return fill_value(v.get("fill_value"), m->dt, &m->fill);
}

std::string chunk_key(const ArrayMeta& m, const std::vector<int64_t>& idx) {
    std::string k = m.key_kind == "default" ? "c" : "";
    if (idx.empty()) return m.key_kind == "default" ? "c" : "0";
    for (size_t d = 0; d < idx.size(); ++d) {
        if (!k.empty()) k += m.sep;
        k += std::to_string(idx[d]);
    }
    return k;
}

ddrs_status read_array(const std::string& dir, const ArrayMeta& m, unsigned char* out) {
    const int esz = m.dt->size;
    const size_t nd = m.shape.size();
    std::vector<int64_t> nchunk(nd);
    int64_t total_chunks = 1;
    for (size_t d = 0; d < nd; ++d) {
        nchunk[d] = (m.shape[d] + m.chunk[d] - 1) / m.chunk[d];
        total_chunks *= nchunk[d];
    }
    int64_t chunk_elems = 1;
    for (int64_t c : m.chunk) chunk_elems *= c;
    const size_t chunk_bytes = static_cast<size_t>(chunk_elems) * esz;
    // the whole array starts as the fill value
    const int64_t n = m.count();
    for (int64_t i = 0; i < n; ++i) std::memcpy(out + i * esz, m.fill.data(), esz);
    std::vector<char> raw;
    std::vector<unsigned char> dec(chunk_bytes);
    std::vector<int64_t> ci(nd, 0), e(nd, 0);
    for (int64_t q = 0; q < total_chunks && n > 0; ++q) {
        int64_t r = q;
        for (size_t d = nd; d-- > 0;) {
            ci[d] = r % nchunk[d];
            r /= nchunk[d];
        }
        const std::string path = join(dir, chunk_key(m, ci));
        if (!read_file(path, &raw)) continue;  // missing chunk: fill value
        const unsigned char* src = reinterpret_cast<const unsigned char*>(raw.data());
        if (m.zstd) {
            const Zstd& z = zstd();
            if (!z.ok) return zfail(DDRS_ERR_INVALID, "zarr: libzstd.so.1 not available");
            const size_t got = z.decompress(dec.data(), dec.size(), raw.data(), raw.size());
            if (z.is_error(got))
                return zfail(DDRS_ERR_INVALID, "zarr: zstd error in " + path + ": " + z.err_name(got));
            if (got != chunk_bytes)
                return zfail(DDRS_ERR_INVALID, "zarr: chunk " + path + " decodes to " +
                                                   std::to_string(got) + " bytes, expected " +
                                                   std::to_string(chunk_bytes));
            src = dec.data();
        } else if (raw.size() != chunk_bytes) {
            return zfail(DDRS_ERR_INVALID, "zarr: chunk " + path + " has " +
                                               std::to_string(raw.size()) + " bytes, expected " +
                                               std::to_string(chunk_bytes));
        }
        // copy the chunk's in-bounds elements (C order), swapping big-endian bytes
        for (int64_t k = 0; k < chunk_elems; ++k) {
            int64_t kk = k, off = 0;
            bool inside = true;
            for (size_t d = nd; d-- > 0;) {
                e[d] = kk % m.chunk[d];
                kk /= m.chunk[d];
            }
            for (size_t d = 0; d < nd; ++d) {
                const int64_t g = ci[d] * m.chunk[d] + e[d];
                if (g >= m.shape[d]) { inside = false; break; }
                off = off * m.shape[d] + g;
            }
            if (!inside) continue;
            unsigned char* dst = out + off * esz;
            const unsigned char* s = src + k * esz;
            if (m.big_endian)
                for (int b = 0; b < esz; ++b) dst[b] = s[esz - 1 - b];
            else
                std::memcpy(dst, s, esz);
        }
    }
    return DDRS_OK;
}

}  // namespace

extern "C" {

ddrs_status ddrs_zarr_array_info(const char* array_dir, int32_t* dtype, int32_t* ndim,
                                 int64_t* shape, int32_t shape_cap, int64_t* count) {
    if (!array_dir) return zfail(DDRS_ERR_INVALID, "zarr: NULL path");
    ArrayMeta m;
    ddrs_status st = read_meta(array_dir, &m);
    if (st != DDRS_OK) return st;
    if (dtype) *dtype = m.dt->code;
    if (ndim) *ndim = static_cast<int32_t>(m.shape.size());
    if (shape)
        for (size_t d = 0; d < m.shape.size() && static_cast<int32_t>(d) < shape_cap; ++d)
            shape[d] = m.shape[d];
    if (count) *count = m.count();
    return DDRS_OK;
}

ddrs_status ddrs_zarr_read_array(const char* array_dir, int32_t dtype, void* h_out,
                                 int64_t capacity) {
    if (!array_dir || (!h_out && capacity > 0)) return zfail(DDRS_ERR_INVALID, "zarr: NULL argument");
    ArrayMeta m;
    ddrs_status st = read_meta(array_dir, &m);
    if (st != DDRS_OK) return st;
    if (m.dt->code != dtype)
        return zfail(DDRS_ERR_INVALID, std::string("zarr: ") + array_dir + " holds " + m.dt->name +
                                           ", not the requested type");
    if (m.count() > capacity)
        return zfail(DDRS_ERR_CAPACITY, std::string("zarr: ") + array_dir + " has " +
                                            std::to_string(m.count()) + " elements, capacity " +
                                            std::to_string(capacity));
    return read_array(array_dir, m, static_cast<unsigned char*>(h_out));
}

ddrs_status ddrs_zarr_group_attr(const char* group_dir, const char* key, double* h_num,
                                 char* h_text, int32_t text_cap) {
    if (!group_dir || !key) return zfail(DDRS_ERR_INVALID, "zarr: NULL argument");
    JVal v;
    ddrs_status st = parse_json_file(join(group_dir, "zarr.json"), &v);
    if (st != DDRS_OK) return st;
    const JVal* nt = v.get("node_type");
    if (!nt || nt->str != "group")
        return zfail(DDRS_ERR_INVALID, std::string("zarr: ") + group_dir + " is not a group");
    const JVal* attrs = v.get("attributes");
    const JVal* a = attrs ? attrs->get(key) : nullptr;
    if (!a) return zfail(DDRS_ERR_INVALID, std::string("zarr: no attribute '") + key + "'");
    if (h_num) *h_num = a->kind == JVal::NUM ? a->num : NAN;
    if (h_text && text_cap > 0) {
        const std::string t = (a->kind == JVal::STR || a->kind == JVal::NUM) ? a->str : "";
        std::snprintf(h_text, static_cast<size_t>(text_cap), "%s", t.c_str());
    }
    return DDRS_OK;
}

}  // extern "C"
