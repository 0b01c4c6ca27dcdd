// capture.cpp -- the sphWater capture format (main.cpp:2438-2552, written by
// common/common-cpp/src/state_writer.cpp:13-87 and read by
// common/common-cpp/src/state_reader.cpp), host code on top of the library's
// C ABI.  A capture is a directory capture_NNNN/ holding state.json
//   {"buffers": [{"bytes", "count", "file", "format", "name", "stride"}, ...],
//    "frame": N, "meta": {"sphWater": {"iteration", "particle_count",
//    "preset_index", "substeps", "tier_index"[, "camera"]}}}
// and one raw <name>.bin per buffer: particles (128-byte records,
// shaders/_struct_layouts.txt:11-42), density_alpha (vec4 rho, alpha/rho^2,
// rho*, rho_dot), pressure, morton_codes, sorted_index, cell_starts.  JSON is
// written the way nlohmann::json::dump(2) writes it (keys sorted, two-space
// indent) and read with a small parser of the full JSON grammar.
#include <sys/stat.h>
#include <dirent.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "dfsph_internal.h"

using namespace dfsph;

namespace {

// ---- JSON ---------------------------------------------------------------------
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    std::string text;                      // number (as written) or string value
    std::vector<Json> arr;
    std::map<std::string, Json> obj;       // sorted keys, as nlohmann::json

    static Json num(uint64_t v) { Json j; j.kind = Number; j.text = std::to_string(v); return j; }
    static Json num(int64_t v) { Json j; j.kind = Number; j.text = std::to_string(v); return j; }
    static Json str(const std::string& s) { Json j; j.kind = String; j.text = s; return j; }
    static Json object() { Json j; j.kind = Object; return j; }
    static Json array() { Json j; j.kind = Array; return j; }
    const Json* get(const std::string& k) const {
        if (kind != Object) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

class Parser {
public:
    explicit Parser(std::string s) : s_(std::move(s)) {}
    bool parse(Json& out, std::string& err) {
        if (!value(out, 0)) { err = err_; return false; }
        ws();
        if (p_ != s_.size()) { err = at("trailing characters"); return false; }
        return true;
    }

private:
    const std::string s_;
    size_t p_ = 0;
    std::string err_;

    std::string at(const char* what) const { return std::string(what) + " at offset " + std::to_string(p_); }
    bool fail(const char* what) { err_ = at(what); return false; }
    void ws() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\r' || s_[p_] == '\n')) ++p_;
    }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if (s_.compare(p_, n, w) != 0) return false;
        p_ += n;
        return true;
    }
    static void utf8(uint32_t cp, std::string& o) {
        if (cp < 0x80) o += char(cp);
        else if (cp < 0x800) { o += char(0xC0 | (cp >> 6)); o += char(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            o += char(0xE0 | (cp >> 12)); o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F));
        } else {
            o += char(0xF0 | (cp >> 18)); o += char(0x80 | ((cp >> 12) & 0x3F));
            o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F));
        }
    }
    // The rest of a multi-byte UTF-8 sequence whose lead byte c was just read:
    // well-formed sequences only (RFC 3629 table: no overlong forms, no
    // surrogates, nothing past U+10FFFF), which is what nlohmann::json accepts.
    bool utf8_tail(unsigned char c, std::string& o) {
        int more;
        unsigned char lo = 0x80, hi = 0xBF;   // range of the first continuation byte
        if (c >= 0xC2 && c <= 0xDF) more = 1;
        else if (c == 0xE0) { more = 2; lo = 0xA0; }
        else if (c == 0xED) { more = 2; hi = 0x9F; }
        else if (c >= 0xE1 && c <= 0xEF) more = 2;
        else if (c == 0xF0) { more = 3; lo = 0x90; }
        else if (c >= 0xF1 && c <= 0xF3) more = 3;
        else if (c == 0xF4) { more = 3; hi = 0x8F; }
        else return fail("ill-formed UTF-8");
        o += char(c);
        for (int q = 0; q < more; ++q) {
            if (p_ >= s_.size()) return fail("ill-formed UTF-8");
            const unsigned char d = static_cast<unsigned char>(s_[p_]);
            if (d < (q == 0 ? lo : 0x80) || d > (q == 0 ? hi : 0xBF)) return fail("ill-formed UTF-8");
            o += char(d);
            ++p_;
        }
        return true;
    }
    bool hex4(uint32_t& v) {
        if (p_ + 4 > s_.size()) return fail("short \\u escape");
        v = 0;
        for (int q = 0; q < 4; ++q) {
            const char c = s_[p_++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool string(std::string& o) {
        if (p_ >= s_.size() || s_[p_] != '"') return fail("expected string");
        ++p_;
        while (true) {
            if (p_ >= s_.size()) return fail("unterminated string");
            const char c = s_[p_++];
            if (c == '"') return true;
            if (static_cast<unsigned char>(c) < 0x20) return fail("control character in string");
            if (static_cast<unsigned char>(c) >= 0x80) {
                if (!utf8_tail(static_cast<unsigned char>(c), o)) return false;
                continue;
            }
            if (c != '\\') { o += c; continue; }
            if (p_ >= s_.size()) return fail("unterminated escape");
            const char e = s_[p_++];
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00) {   // surrogate pair
                        uint32_t lo;
                        if (!lit("\\u") || !hex4(lo) || lo < 0xDC00 || lo >= 0xE000)
                            return fail("bad surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        return fail("low surrogate without a high one");
                    }
                    utf8(cp, o);
                    break;
                }
                default: return fail("bad escape");
            }
        }
    }
    bool number(Json& out) {
        const size_t st = p_;
        if (p_ < s_.size() && s_[p_] == '-') ++p_;
        auto digits = [&] { const size_t a = p_; while (p_ < s_.size() && std::isdigit((unsigned char)s_[p_])) ++p_; return p_ > a; };
        if (p_ < s_.size() && s_[p_] == '0') ++p_;
        else if (!digits()) return fail("bad number");
        if (p_ < s_.size() && s_[p_] == '.') { ++p_; if (!digits()) return fail("bad fraction"); }
        if (p_ < s_.size() && (s_[p_] == 'e' || s_[p_] == 'E')) {
            ++p_;
            if (p_ < s_.size() && (s_[p_] == '+' || s_[p_] == '-')) ++p_;
            if (!digits()) return fail("bad exponent");
        }
        out.kind = Json::Number;
        out.text = s_.substr(st, p_ - st);
        return true;
    }
    bool value(Json& out, int depth) {
        if (depth > 256) return fail("nesting too deep");
        ws();
        if (p_ >= s_.size()) return fail("unexpected end");
        const char c = s_[p_];
        if (c == '{') {
            ++p_;
            out.kind = Json::Object;
            ws();
            if (p_ < s_.size() && s_[p_] == '}') { ++p_; return true; }
            while (true) {
                ws();
                std::string k;
                if (!string(k)) return false;
                ws();
                if (p_ >= s_.size() || s_[p_] != ':') return fail("expected ':'");
                ++p_;
                Json v;
                if (!value(v, depth + 1)) return false;
                out.obj[k] = std::move(v);
                ws();
                if (p_ < s_.size() && s_[p_] == ',') { ++p_; continue; }
                if (p_ < s_.size() && s_[p_] == '}') { ++p_; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++p_;
            out.kind = Json::Array;
            ws();
            if (p_ < s_.size() && s_[p_] == ']') { ++p_; return true; }
            while (true) {
                Json v;
                if (!value(v, depth + 1)) return false;
                out.arr.push_back(std::move(v));
                ws();
                if (p_ < s_.size() && s_[p_] == ',') { ++p_; continue; }
                if (p_ < s_.size() && s_[p_] == ']') { ++p_; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') { out.kind = Json::String; return string(out.text); }
        if (lit("true")) { out.kind = Json::Bool; out.b = true; return true; }
        if (lit("false")) { out.kind = Json::Bool; out.b = false; return true; }
        if (lit("null")) { out.kind = Json::Null; return true; }
        return number(out);
    }
};

void escape(const std::string& s, std::string& o) {
    o += '"';
    for (const char ch : s) {
        const unsigned char c = static_cast<unsigned char>(ch);
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\b': o += "\\b"; break;
            case '\f': o += "\\f"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            default:
                if (c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof(buf), "\\u%04x", c);
                    o += buf;
                } else {
                    o += ch;
                }
        }
    }
    o += '"';
}

// nlohmann::json::dump(indent) layout.
void dump(const Json& j, int indent, int level, std::string& o) {
    const std::string pad(size_t(indent) * size_t(level + 1), ' ');
    const std::string close(size_t(indent) * size_t(level), ' ');
    switch (j.kind) {
        case Json::Null: o += "null"; break;
        case Json::Bool: o += j.b ? "true" : "false"; break;
        case Json::Number: o += j.text; break;
        case Json::String: escape(j.text, o); break;
        case Json::Array:
            if (j.arr.empty()) { o += "[]"; break; }
            o += "[\n";
            for (size_t q = 0; q < j.arr.size(); ++q) {
                o += pad;
                dump(j.arr[q], indent, level + 1, o);
                o += q + 1 < j.arr.size() ? ",\n" : "\n";
            }
            o += close + "]";
            break;
        case Json::Object: {
            if (j.obj.empty()) { o += "{}"; break; }
            o += "{\n";
            size_t q = 0;
            for (const auto& kv : j.obj) {
                o += pad;
                escape(kv.first, o);
                o += ": ";
                dump(kv.second, indent, level + 1, o);
                o += ++q < j.obj.size() ? ",\n" : "\n";
            }
            o += close + "}";
            break;
        }
    }
}

// Unsigned integer value of a JSON number, or false.
bool as_u64(const Json* j, uint64_t& v) {
    if (!j || j->kind != Json::Number || j->text.empty() || j->text[0] == '-') return false;
    for (const char c : j->text)
        if (!std::isdigit(static_cast<unsigned char>(c))) return false;
    errno = 0;
    v = std::strtoull(j->text.c_str(), nullptr, 10);
    return errno == 0;
}
bool as_i64(const Json* j, int64_t& v) {
    if (!j || j->kind != Json::Number || j->text.empty()) return false;
    for (size_t q = j->text[0] == '-' ? 1 : 0; q < j->text.size(); ++q)
        if (!std::isdigit(static_cast<unsigned char>(j->text[q]))) return false;
    errno = 0;
    v = std::strtoll(j->text.c_str(), nullptr, 10);
    return errno == 0;
}

// ---- files ----------------------------------------------------------------------
bool make_dirs(const std::string& path) {
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part;
        if (::mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) return false;
        cur += '/';
    }
    struct stat st{};
    return ::stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool write_file(const std::string& path, const void* data, size_t bytes) {
    std::ofstream out(path, std::ios::binary);
    if (!out) return false;
    out.write(static_cast<const char*>(data), std::streamsize(bytes));
    return out.good();
}

bool read_file(const std::string& path, std::vector<uint8_t>& bytes) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) return false;
    const std::streamsize size = in.tellg();
    if (size < 0) return false;
    in.seekg(0);
    bytes.resize(size_t(size));
    return size == 0 || bool(in.read(reinterpret_cast<char*>(bytes.data()), size));
}

std::string frame_dir(uint32_t frame) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "capture_%04u", frame);
    return buf;
}

int copy_out(const std::string& s, char* dst, size_t bytes) {
    if (!dst) return DFSPH_OK;
    if (s.size() + 1 > bytes) return set_error(DFSPH_E_INVALID, "output path buffer too small (" +
                                                                      std::to_string(s.size() + 1) + " bytes needed)");
    std::memcpy(dst, s.c_str(), s.size() + 1);
    return DFSPH_OK;
}

// ---- reading a capture ------------------------------------------------------------
constexpr size_t kRecord = 128;   // particles_buf record (shaders/_struct_layouts.txt:11-42)

struct Capture {
    Json state;
    const Json* meta = nullptr;   // state.meta.sphWater
    uint64_t count = 0;
    std::map<std::string, std::vector<uint8_t>> bufs;
};

// The checks of capture_load (main.cpp:2493-2527): the sphWater meta, the
// saved particle count, and the five per-particle buffers sized from it.
int open_capture(const std::string& dir, bool load_buffers, Capture& cap) {
    std::vector<uint8_t> text;
    if (!read_file(dir + "/state.json", text))
        return set_error(DFSPH_E_IO, "cannot open " + dir + "/state.json");
    std::string err;
    Parser ps(std::string(text.begin(), text.end()));
    if (!ps.parse(cap.state, err)) return set_error(DFSPH_E_IO, "invalid JSON in state.json: " + err);
    const Json* meta = cap.state.get("meta");
    cap.meta = meta ? meta->get("sphWater") : nullptr;
    if (!cap.meta || cap.meta->kind != Json::Object)
        return set_error(DFSPH_E_IO, "capture has no sphWater meta");
    if (!as_u64(cap.meta->get("particle_count"), cap.count) || cap.count > 0xFFFFFFFFull)
        return set_error(DFSPH_E_IO, "sphWater.particle_count missing or not a 32-bit count");
    const Json* bl = cap.state.get("buffers");
    if (!bl || bl->kind != Json::Array) return set_error(DFSPH_E_IO, "state.json has no buffers array");
    struct Need { const char* name; size_t stride; bool required; };
    const Need need[] = {{"particles", kRecord, true},   {"density_alpha", 16, true}, {"pressure", 4, true},
                         {"morton_codes", 4, true},     {"sorted_index", 4, true},   {"cell_starts", 4, false}};
    for (const Need& nd : need) {
        const Json* desc = nullptr;
        for (const Json& b : bl->arr) {
            const Json* nm = b.get("name");
            if (nm && nm->kind == Json::String && nm->text == nd.name) { desc = &b; break; }
        }
        if (!desc) {
            if (nd.required) return set_error(DFSPH_E_IO, std::string("capture has no buffer ") + nd.name);
            continue;
        }
        const Json* file = desc->get("file");
        const std::string fname = file && file->kind == Json::String ? file->text : std::string(nd.name) + ".bin";
        if (fname.find('/') != std::string::npos || fname == "..")
            return set_error(DFSPH_E_IO, "buffer file name " + fname + " leaves the capture directory");
        std::vector<uint8_t> bytes;
        if (!read_file(dir + "/" + fname, bytes))
            return set_error(DFSPH_E_IO, "cannot read " + dir + "/" + fname);
        if (nd.required && bytes.size() != size_t(cap.count) * nd.stride)
            return set_error(DFSPH_E_IO, std::string("buffer-size mismatch: ") + nd.name + " holds " +
                                                  std::to_string(bytes.size()) + " bytes for particle_count " +
                                                  std::to_string(cap.count));
        if (load_buffers) cap.bufs[nd.name] = std::move(bytes);
    }
    return DFSPH_OK;
}

void fill_info(const Capture& cap, dfsph_capture_info* info) {
    if (!info) return;
    uint64_t it = 0;
    int64_t v = 0;
    info->iteration = as_u64(cap.meta->get("iteration"), it) ? it : 0;
    info->particle_count = uint32_t(cap.count);
    info->tier_index = as_i64(cap.meta->get("tier_index"), v) ? int32_t(v) : 0;
    info->preset_index = as_i64(cap.meta->get("preset_index"), v) ? int32_t(v) : 0;
    info->substeps = as_i64(cap.meta->get("substeps"), v) ? int32_t(v) : 0;
    uint64_t frame = 0;
    info->frame = as_u64(cap.state.get("frame"), frame) ? uint32_t(frame) : 0u;
}

}  // namespace

extern "C" {

int dfsph_capture_save(dfsph_ctx* c, const char* root_dir, uint32_t frame, const dfsph_capture_info* info,
                       char* dir_out, size_t dir_bytes) {
    if (!c || !root_dir) return set_error(DFSPH_E_INVALID, "NULL argument");
    size_t pos_bytes = 0;
    if (int rc = dfsph_field_size(c, DFSPH_FIELD_POSITION, &pos_bytes)) return rc;
    const size_t n = pos_bytes / 12;
    if (n == 0) return set_error(DFSPH_E_INVALID, "no particles to capture");   // main.cpp:2441-2444
    Json camera;
    if (info && info->camera_json) {
        std::string err;
        Parser ps{std::string(info->camera_json)};
        if (!ps.parse(camera, err) || camera.kind != Json::Object)
            return set_error(DFSPH_E_INVALID, "camera_json must be a JSON object: " + err);
    }
    // Fields in original-index order, the order of the reference's buffers.
    std::vector<float> x(n * 3), v(n * 3), rho(n, 0.0f), aor2(n, 0.0f), rho_adv(n, 0.0f), rho_dot(n, 0.0f),
        p(n, 0.0f);
    std::vector<uint32_t> keys(n, 0u), sorted(n, 0u), starts;
    int rc;
    if ((rc = dfsph_export(c, DFSPH_FIELD_POSITION, x.data(), n * 12)) ||
        (rc = dfsph_export(c, DFSPH_FIELD_VELOCITY, v.data(), n * 12)))
        return rc;
    const bool stepped = dfsph_export(c, DFSPH_FIELD_DENSITY, rho.data(), n * 4) == DFSPH_OK;
    if (stepped) {
        size_t cs = 0;
        if ((rc = dfsph_export(c, DFSPH_FIELD_ALPHA_OVER_RHO2, aor2.data(), n * 4)) ||
            (rc = dfsph_export(c, DFSPH_FIELD_DENSITY_ADV, rho_adv.data(), n * 4)) ||
            (rc = dfsph_export(c, DFSPH_FIELD_DENSITY_CHANGE, rho_dot.data(), n * 4)) ||
            (rc = dfsph_export(c, DFSPH_FIELD_PRESSURE, p.data(), n * 4)) ||
            (rc = dfsph_export(c, DFSPH_FIELD_MORTON_KEY, keys.data(), n * 4)) ||
            (rc = dfsph_export(c, DFSPH_FIELD_SORTED_INDEX, sorted.data(), n * 4)) ||
            (rc = dfsph_field_size(c, DFSPH_FIELD_CELL_STARTS, &cs)))
            return rc;
        starts.resize(cs / 4);
        if ((rc = dfsph_export(c, DFSPH_FIELD_CELL_STARTS, starts.data(), cs))) return rc;
    } else {
        set_error(DFSPH_OK, "");
    }
    const float radius = ctx_particle_radius(c);
    std::vector<uint8_t> rec(n * kRecord, 0);
    std::vector<float> da(n * 4);
    for (size_t i = 0; i < n; ++i) {
        float* r = reinterpret_cast<float*>(rec.data() + i * kRecord);
        r[0] = x[3 * i]; r[1] = x[3 * i + 1]; r[2] = x[3 * i + 2]; r[3] = radius;   // slot 0
        r[4] = v[3 * i]; r[5] = v[3 * i + 1]; r[6] = v[3 * i + 2];                  // slot 1
        const uint32_t id = uint32_t(i), flags = ctx_particle_flags(c, uint32_t(i));
        std::memcpy(r + 20, &id, 4);                                                 // slot 5.x
        std::memcpy(r + 22, &flags, 4);                                              // slot 5.z
        da[4 * i] = rho[i]; da[4 * i + 1] = aor2[i]; da[4 * i + 2] = rho_adv[i]; da[4 * i + 3] = rho_dot[i];
    }
    const std::string dir = std::string(root_dir) + "/" + frame_dir(frame);
    if (!make_dirs(dir)) return set_error(DFSPH_E_IO, "cannot create " + dir);
    Json state = Json::object();
    state.obj["frame"] = Json::num(uint64_t(frame));
    Json meta = Json::object();
    meta.obj["iteration"] = Json::num(uint64_t(info ? info->iteration : 0));
    meta.obj["particle_count"] = Json::num(uint64_t(n));
    meta.obj["tier_index"] = Json::num(int64_t(info ? info->tier_index : 0));
    meta.obj["preset_index"] = Json::num(int64_t(info ? info->preset_index : 0));
    meta.obj["substeps"] = Json::num(int64_t(info ? info->substeps : 1));
    if (camera.kind == Json::Object) meta.obj["camera"] = camera;
    state.obj["meta"] = Json::object();
    state.obj["meta"].obj["sphWater"] = meta;
    state.obj["buffers"] = Json::array();
    auto save = [&](const char* name, const void* data, size_t bytes, uint64_t count, uint64_t stride) -> int {
        const std::string file = std::string(name) + ".bin";
        if (!write_file(dir + "/" + file, data, bytes))
            return set_error(DFSPH_E_IO, "cannot write " + dir + "/" + file);
        Json d = Json::object();
        d.obj["count"] = Json::num(count);
        d.obj["stride"] = Json::num(stride);
        d.obj["format"] = Json::str("raw");
        d.obj["name"] = Json::str(name);
        d.obj["file"] = Json::str(file);
        d.obj["bytes"] = Json::num(uint64_t(bytes));
        state.obj["buffers"].arr.push_back(d);
        return DFSPH_OK;
    };
    if ((rc = save("particles", rec.data(), rec.size(), n, kRecord)) ||
        (rc = save("density_alpha", da.data(), da.size() * 4, n, 16)) ||
        (rc = save("pressure", p.data(), n * 4, n, 4)) ||
        (rc = save("morton_codes", keys.data(), n * 4, n, 4)) ||
        (rc = save("sorted_index", sorted.data(), n * 4, n, 4)) ||
        (rc = save("cell_starts", starts.data(), starts.size() * 4, starts.size(), 4)))
        return rc;
    std::string text;
    dump(state, 2, 0, text);
    if (!write_file(dir + "/state.json", text.data(), text.size()))
        return set_error(DFSPH_E_IO, "cannot write " + dir + "/state.json");
    return copy_out(dir, dir_out, dir_bytes);
}

int dfsph_capture_read_info(const char* capture_dir, dfsph_capture_info* info) {
    if (!capture_dir || !info) return set_error(DFSPH_E_INVALID, "NULL argument");
    Capture cap;
    if (int rc = open_capture(capture_dir, false, cap)) return rc;
    fill_info(cap, info);
    return DFSPH_OK;
}

int dfsph_capture_load(dfsph_ctx* c, const char* capture_dir, int32_t boundary, uint32_t boundary_cap,
                       dfsph_capture_info* info) {
    if (!c || !capture_dir) return set_error(DFSPH_E_INVALID, "NULL argument");
    Capture cap;
    if (int rc = open_capture(capture_dir, true, cap)) return rc;
    const size_t n = size_t(cap.count);
    // The reference loads an empty capture as an empty tier; a scene here
    // holds at least one particle, as capture_save requires.
    if (n == 0) return set_error(DFSPH_E_INVALID, "capture holds no particles");
    const std::vector<uint8_t>& rec = cap.bufs["particles"];
    std::vector<float> x(n * 3), v(n * 3), p(n);
    std::vector<uint32_t> flags(n);
    for (size_t i = 0; i < n; ++i) {
        const float* r = reinterpret_cast<const float*>(rec.data() + i * kRecord);
        for (int a = 0; a < 3; ++a) {
            x[3 * i + a] = r[a];        // slot 0
            v[3 * i + a] = r[4 + a];    // slot 1
        }
        std::memcpy(&flags[i], r + 22, 4);
    }
    std::memcpy(p.data(), cap.bufs["pressure"].data(), n * 4);
    dfsph_scene s{};
    s.count = uint32_t(n);
    s.positions = x.data();
    s.velocities = v.data();
    s.boundary = boundary;
    s.boundary_cap = boundary_cap;
    if (int rc = dfsph_init_scene(c, &s)) return rc;
    uint64_t it = 0;
    as_u64(cap.meta->get("iteration"), it);
    if (int rc = ctx_resume(c, p.data(), flags.data(), it)) return rc;
    fill_info(cap, info);
    return DFSPH_OK;
}

int dfsph_capture_find_latest(const char* root_dir, char* dir_out, size_t dir_bytes) {
    if (!root_dir || !dir_out) return set_error(DFSPH_E_INVALID, "NULL argument");
    // StateReader::findLatest (state_reader.cpp:12-28): the last capture_*
    // directory in name order.
    DIR* d = ::opendir(root_dir);
    if (!d) return set_error(DFSPH_E_IO, std::string("cannot open directory ") + root_dir);
    std::vector<std::string> names;
    while (dirent* e = ::readdir(d)) {
        const std::string nm = e->d_name;
        if (nm.rfind("capture_", 0) != 0) continue;
        struct stat st{};
        const std::string full = std::string(root_dir) + "/" + nm;
        if (::stat(full.c_str(), &st) == 0 && S_ISDIR(st.st_mode)) names.push_back(nm);
    }
    ::closedir(d);
    if (names.empty()) return set_error(DFSPH_E_IO, std::string("no captures in ") + root_dir);
    std::sort(names.begin(), names.end());
    return copy_out(std::string(root_dir) + "/" + names.back(), dir_out, dir_bytes);
}

}  // extern "C"
