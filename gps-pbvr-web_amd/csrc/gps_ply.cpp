// 3DGS PLY ingest: the on-disk format in front of the renderer (crates/gps-io/src/lib.rs:1-373) and the
// activation to GpuGaussian (crates/gps-render/src/gps.rs:41-49).  Host code only.
//
// Accepted: `format ascii` or `format binary_little_endian`, one `vertex` element first, scalar properties
// of any PLY type (converted through f64 to f32, as gps-io does), SH degree 0..3 from 0/9/24/45 `f_rest_*`
// properties; other properties (normals, ...) are skipped.  Values stay in file form (log-scales, opacity
// logit, (w, x, y, z) rotation, f_dc) until gps_ply_to_gpu applies the activation.  Every failure returns an
// error code with the message gps-io's PlyError would print; nothing aborts on bad input.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/gps_hip.h"

struct GpsPly {
    uint64_t count = 0;
    uint32_t sh_degree = 0;
    int binary = 0;
    std::vector<float> position, log_scale, rot_wxyz, opacity_logit, f_dc, sh_rest;
};

namespace {

thread_local std::string g_ply_error;

enum Scalar { I8, U8, I16, U16, I32, U32, F32, F64, BAD };

Scalar scalar_of(const std::string &s) {
    if (s == "char" || s == "int8") return I8;
    if (s == "uchar" || s == "uint8") return U8;
    if (s == "short" || s == "int16") return I16;
    if (s == "ushort" || s == "uint16") return U16;
    if (s == "int" || s == "int32") return I32;
    if (s == "uint" || s == "uint32") return U32;
    if (s == "float" || s == "float32") return F32;
    if (s == "double" || s == "float64") return F64;
    return BAD;
}

size_t size_of(Scalar t) {
    switch (t) {
    case I8: case U8: return 1;
    case I16: case U16: return 2;
    case F64: return 8;
    default: return 4;
    }
}

// little-endian read of one scalar, widened to f64 (gps-io read_le)
double read_le(Scalar t, const uint8_t *b) {
    uint64_t u = 0;
    for (size_t k = 0; k < size_of(t); k++) u |= (uint64_t)b[k] << (8 * k);
    switch (t) {
    case I8: return (double)(int8_t)u;
    case U8: return (double)(uint8_t)u;
    case I16: return (double)(int16_t)u;
    case U16: return (double)(uint16_t)u;
    case I32: return (double)(int32_t)u;
    case U32: return (double)(uint32_t)u;
    case F32: {
        uint32_t v = (uint32_t)u;
        float f;
        std::memcpy(&f, &v, 4);
        return (double)f;
    }
    default: {
        double d;
        std::memcpy(&d, &u, 8);
        return d;
    }
    }
}

struct Prop {
    std::string name;
    Scalar type;
    size_t offset;
};

struct Header {
    bool binary = false;
    uint64_t vertex_count = 0;
    std::vector<Prop> props;
    size_t stride = 0;
    size_t data_start = 0;
};

int fail(int code, const std::string &msg) {
    g_ply_error = msg;
    return code;
}

std::vector<std::string> split_ws(const std::string &line) {
    std::vector<std::string> t;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && std::isspace((unsigned char)line[i])) i++;
        size_t j = i;
        while (j < line.size() && !std::isspace((unsigned char)line[j])) j++;
        if (j > i) t.push_back(line.substr(i, j - i));
        i = j;
    }
    return t;
}

// gps-io parse_header (lib.rs:148-227)
int parse_header(const uint8_t *bytes, size_t len, Header &h) {
    static const char marker[] = "end_header";
    const size_t ml = sizeof(marker) - 1;
    size_t pos = SIZE_MAX;
    for (size_t i = 0; i + ml <= len; i++)
        if (std::memcmp(bytes + i, marker, ml) == 0) {
            pos = i;
            break;
        }
    if (pos == SIZE_MAX) return fail(GPS_PLY_BAD_HEADER, "bad PLY header: no end_header");
    size_t ds = pos + ml;
    if (ds < len && bytes[ds] == '\r') ds++;
    if (ds < len && bytes[ds] == '\n') ds++;
    else return fail(GPS_PLY_BAD_HEADER, "bad PLY header: end_header not followed by newline");
    for (size_t i = 0; i < pos; i++)
        if (bytes[i] >= 0x80) return fail(GPS_PLY_BAD_HEADER, "bad PLY header: non-UTF8 header");
    const std::string text((const char *)bytes, pos);
    std::vector<std::string> lines;
    {
        size_t i = 0;
        while (i <= text.size()) {
            size_t j = text.find('\n', i);
            if (j == std::string::npos) j = text.size();
            std::string l = text.substr(i, j - i);
            if (!l.empty() && l.back() == '\r') l.pop_back();
            if (j < text.size() || !l.empty()) lines.push_back(l);
            i = j + 1;
        }
    }
    auto trim = [](const std::string &s) {
        size_t a = 0, b = s.size();
        while (a < b && std::isspace((unsigned char)s[a])) a++;
        while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
        return s.substr(a, b - a);
    };
    if (lines.empty() || trim(lines[0]) != "ply") return fail(GPS_PLY_NOT_PLY, "not a PLY file");
    int binary = -1;
    bool have_count = false, in_vertex = false, other_before = false;
    size_t offset = 0;
    for (size_t li = 1; li < lines.size(); li++) {
        const std::vector<std::string> t = split_ws(lines[li]);
        if (t.size() == 3 && t[0] == "format") {
            if (t[1] == "ascii") binary = 0;
            else if (t[1] == "binary_little_endian") binary = 1;
            else return fail(GPS_PLY_UNSUPPORTED, "unsupported PLY format: " + t[1]);
        } else if (t.size() == 3 && t[0] == "element") {
            in_vertex = t[1] == "vertex";
            if (in_vertex) {
                errno = 0;
                char *end = nullptr;
                const std::string &ns = t[2];
                unsigned long long n = std::strtoull(ns.c_str(), &end, 10);
                if (ns.empty() || ns[0] == '-' || *end != '\0' || errno == ERANGE)
                    return fail(GPS_PLY_BAD_HEADER, "bad PLY header: vertex count");
                if (n > GPS_PLY_MAX_VERTICES)
                    return fail(GPS_PLY_TOO_MANY_VERTICES, "vertex count " + std::to_string(n) + " exceeds limit " +
                                                               std::to_string((unsigned long long)GPS_PLY_MAX_VERTICES));
                h.vertex_count = n;
                have_count = true;
            } else if (!have_count) {
                other_before = true;   // an element before "vertex" would shift the data
            }
        } else if (t.size() >= 2 && t[0] == "property" && t[1] == "list") {
            if (in_vertex) return fail(GPS_PLY_UNSUPPORTED, "unsupported PLY format: list property in vertex element");
        } else if (t.size() == 3 && t[0] == "property" && in_vertex) {
            const Scalar s = scalar_of(t[1]);
            if (s == BAD) return fail(GPS_PLY_BAD_HEADER, "bad PLY header: type " + t[1]);
            h.props.push_back({t[2], s, offset});
            offset += size_of(s);
        }
    }
    if (other_before) return fail(GPS_PLY_UNSUPPORTED, "unsupported PLY format: element before vertex");
    if (binary < 0) return fail(GPS_PLY_BAD_HEADER, "bad PLY header: no format line");
    if (!have_count) return fail(GPS_PLY_BAD_HEADER, "bad PLY header: no vertex element");
    h.binary = binary == 1;
    h.stride = offset;
    h.data_start = ds;
    return 0;
}

// Rust's str::parse::<f32>: decimal only (no hex), the whole token, correctly rounded
bool parse_f32(const std::string &tok, float &out) {
    if (tok.empty()) return false;
    for (char ch : tok)
        if (ch == 'x' || ch == 'X' || ch == 'p' || ch == 'P') return false;
    errno = 0;
    char *end = nullptr;
    const float v = std::strtof(tok.c_str(), &end);
    if (end != tok.c_str() + tok.size()) return false;
    out = v;   // overflow gives +-inf, as Rust's parse does
    return true;
}

}  // namespace

extern "C" {

const char *gps_ply_last_error(void) { return g_ply_error.c_str(); }

int gps_ply_load(const uint8_t *bytes, uint64_t len, GpsPly **out) {
    *out = nullptr;
    g_ply_error.clear();
    if (!bytes && len) return fail(GPS_PLY_NOT_PLY, "not a PLY file");
    Header h;
    int rc = parse_header(bytes, (size_t)len, h);
    if (rc) return rc;
    auto idx = [&](const char *name, int &at) -> int {
        for (size_t k = 0; k < h.props.size(); k++)
            if (h.props[k].name == name) {
                at = (int)k;
                return 0;
            }
        return fail(GPS_PLY_MISSING_PROPERTY, std::string("missing required property '") + name + "'");
    };
    int ix[3], idc[3], iop, isc[3], irot[4];
    const char *nx[3] = {"x", "y", "z"}, *ndc[3] = {"f_dc_0", "f_dc_1", "f_dc_2"};
    const char *nsc[3] = {"scale_0", "scale_1", "scale_2"}, *nrot[4] = {"rot_0", "rot_1", "rot_2", "rot_3"};
    for (int k = 0; k < 3; k++)
        if ((rc = idx(nx[k], ix[k]))) return rc;
    for (int k = 0; k < 3; k++)
        if ((rc = idx(ndc[k], idc[k]))) return rc;
    if ((rc = idx("opacity", iop))) return rc;
    for (int k = 0; k < 3; k++)
        if ((rc = idx(nsc[k], isc[k]))) return rc;
    for (int k = 0; k < 4; k++)
        if ((rc = idx(nrot[k], irot[k]))) return rc;
    std::vector<int> irest;
    for (;;) {
        const std::string want = "f_rest_" + std::to_string(irest.size());
        int at = -1;
        for (size_t k = 0; k < h.props.size(); k++)
            if (h.props[k].name == want) {
                at = (int)k;
                break;
            }
        if (at < 0) break;
        irest.push_back(at);
    }
    uint32_t deg;
    switch (irest.size()) {
    case 0: deg = 0; break;
    case 9: deg = 1; break;
    case 24: deg = 2; break;
    case 45: deg = 3; break;
    default: return fail(GPS_PLY_BAD_SH_COUNT, std::to_string(irest.size()) + " f_rest_* properties is not 9, 24 or 45");
    }
    const uint64_t n = h.vertex_count;
    const uint8_t *body = bytes + h.data_start;
    const uint64_t blen = len - h.data_start;
    // validate the body length before allocating (lib.rs:279-294)
    if (h.binary) {
        const unsigned __int128 need = (unsigned __int128)n * h.stride;
        if (need > (unsigned __int128)UINT64_MAX) return fail(GPS_PLY_TOO_MANY_VERTICES, "vertex count overflow");
        if (blen < (uint64_t)need)
            return fail(GPS_PLY_TRUNCATED, "file truncated: need " + std::to_string((uint64_t)need) +
                                                " bytes of vertex data, have " + std::to_string(blen));
    } else {
        const uint64_t per = (h.props.empty() ? 1 : h.props.size()) * 2;
        const uint64_t need = n * per;
        if (blen + 1 < need)
            return fail(GPS_PLY_TRUNCATED, "file truncated: need " + std::to_string(need - 1) +
                                                " bytes of vertex data, have " + std::to_string(blen));
    }
    GpsPly *g = new (std::nothrow) GpsPly();
    if (!g) return fail(GPS_PLY_TOO_MANY_VERTICES, "out of host memory");
    const uint32_t per_ch = deg ? (deg + 1) * (deg + 1) - 1 : 0;
    try {
        g->position.resize(3 * n);
        g->log_scale.resize(3 * n);
        g->rot_wxyz.resize(4 * n);
        g->opacity_logit.resize(n);
        g->f_dc.resize(3 * n);
        g->sh_rest.resize((size_t)n * 3 * per_ch);
    } catch (...) {
        delete g;
        return fail(GPS_PLY_TOO_MANY_VERTICES, "vertex count " + std::to_string(n) + " exceeds host memory");
    }
    g->count = n;
    g->sh_degree = deg;
    g->binary = h.binary ? 1 : 0;
    std::vector<float> vals(h.props.size());
    // push_vertex (lib.rs:337-371): finiteness checked per field, in the file's field order
    auto store = [&](uint64_t v) -> int {
        auto chk = [&](float x, const char *field, float &dst) -> int {
            if (!std::isfinite(x))
                return fail(GPS_PLY_NON_FINITE,
                            std::string("non-finite value in '") + field + "' at vertex " + std::to_string(v));
            dst = x;
            return 0;
        };
        int e;
        for (int k = 0; k < 3; k++)
            if ((e = chk(vals[ix[k]], "position", g->position[3 * v + k]))) return e;
        for (int k = 0; k < 3; k++)
            if ((e = chk(vals[idc[k]], "f_dc", g->f_dc[3 * v + k]))) return e;
        if ((e = chk(vals[iop], "opacity", g->opacity_logit[v]))) return e;
        for (int k = 0; k < 3; k++)
            if ((e = chk(vals[isc[k]], "scale", g->log_scale[3 * v + k]))) return e;
        for (int k = 0; k < 4; k++)
            if ((e = chk(vals[irot[k]], "rot", g->rot_wxyz[4 * v + k]))) return e;
        for (size_t k = 0; k < irest.size(); k++)
            if ((e = chk(vals[irest[k]], "f_rest", g->sh_rest[(size_t)v * irest.size() + k]))) return e;
        return 0;
    };
    if (h.binary) {
        for (uint64_t v = 0; v < n; v++) {
            const uint8_t *rec = body + v * h.stride;
            for (size_t k = 0; k < h.props.size(); k++)
                vals[k] = (float)read_le(h.props[k].type, rec + h.props[k].offset);
            if ((rc = store(v))) {
                delete g;
                return rc;
            }
        }
    } else {
        for (uint64_t i = 0; i < blen; i++)
            if (body[i] >= 0x80) {
                delete g;
                return fail(GPS_PLY_BAD_ASCII, "ASCII parse error at vertex 0: non-UTF8 body");
            }
        size_t p = 0;
        auto next_line = [&](std::string &line) -> bool {   // non-blank lines only
            while (p < blen) {
                size_t q = p;
                while (q < blen && body[q] != '\n') q++;
                line.assign((const char *)body + p, q - p);
                p = q < blen ? q + 1 : q;
                if (!split_ws(line).empty()) return true;
            }
            return false;
        };
        std::string line;
        for (uint64_t v = 0; v < n; v++) {
            if (!next_line(line)) {
                delete g;
                return fail(GPS_PLY_TRUNCATED, "file truncated: need " + std::to_string(n) +
                                                    " bytes of vertex data, have " + std::to_string(v));
            }
            const std::vector<std::string> tok = split_ws(line);
            for (size_t k = 0; k < vals.size(); k++) {
                if (k >= tok.size()) {
                    delete g;
                    return fail(GPS_PLY_BAD_ASCII, "ASCII parse error at vertex " + std::to_string(v) + ": expected " +
                                                       std::to_string(vals.size()) + " values, got " + std::to_string(k));
                }
                if (!parse_f32(tok[k], vals[k])) {
                    delete g;
                    return fail(GPS_PLY_BAD_ASCII, "ASCII parse error at vertex " + std::to_string(v) +
                                                       ": invalid float literal");
                }
            }
            if ((rc = store(v))) {
                delete g;
                return rc;
            }
        }
    }
    *out = g;
    return 0;
}

void gps_ply_free(GpsPly *g) { delete g; }

void gps_ply_info(const GpsPly *g, uint64_t *count, uint32_t *sh_degree, int *binary) {
    if (count) *count = g->count;
    if (sh_degree) *sh_degree = g->sh_degree;
    if (binary) *binary = g->binary;
}

void gps_ply_read(const GpsPly *g, float *position, float *log_scale, float *rot_wxyz, float *opacity_logit,
                  float *f_dc, float *sh_rest) {
    auto cp = [](float *dst, const std::vector<float> &src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(float));
    };
    cp(position, g->position);
    cp(log_scale, g->log_scale);
    cp(rot_wxyz, g->rot_wxyz);
    cp(opacity_logit, g->opacity_logit);
    cp(f_dc, g->f_dc);
    cp(sh_rest, g->sh_rest);
}

// Gaussians::activated + GpuGaussian::from_activated (lib.rs:84-96, gps.rs:41-49): f64 sigmoid of the logit,
// colour = f_dc * C0 + 0.5, everything else as stored (the rotation is not normalised)
void gps_ply_to_gpu(const GpsPly *g, GpsGaussian *out) {
    const double C0 = 0.28209479177387814;
    for (uint64_t i = 0; i < g->count; i++) {
        GpsGaussian &o = out[i];
        const float *p = &g->position[3 * i], *s = &g->log_scale[3 * i], *q = &g->rot_wxyz[4 * i],
                    *d = &g->f_dc[3 * i];
        const double op = 1.0 / (1.0 + std::exp(-(double)g->opacity_logit[i]));
        o.pos_opacity[0] = p[0];
        o.pos_opacity[1] = p[1];
        o.pos_opacity[2] = p[2];
        o.pos_opacity[3] = (float)op;
        o.log_scale[0] = s[0];
        o.log_scale[1] = s[1];
        o.log_scale[2] = s[2];
        o.log_scale[3] = 0.0f;
        for (int k = 0; k < 4; k++) o.rot_wxyz[k] = q[k];
        for (int k = 0; k < 3; k++) o.color[k] = (float)((double)d[k] * C0 + 0.5);
        o.color[3] = 0.0f;
    }
}

}  // extern "C"
