// atmo_textures.cpp — the ork texture descriptors through which Proland's
// renderers load the precomputed tables (include/atmo_consumer.h:
// atmo_texture_load_desc, atmo_texture_load).
//
// A descriptor is one XML element (src/demo/textures/atmo/inscatter.xml:25,
// src/atmo/examples/atmo/helloworld.xml:71-75):
//   <texture3D name="inscatter" source="inscatter.raw" internalformat="RGBA16F"
//              min="LINEAR" mag="LINEAR" wraps="CLAMP_TO_EDGE" .../>
// Its source is a .raw file with the 20-byte trailer
// {0xCAFEBABE, width, height * depth, depth, components} that
// PreprocessAtmo.cpp:355-405 writes; a texture2D has depth 0.  The GPU stores
// the payload in the internal format: a 16F format holds every value as the
// nearest IEEE binary16 value, which is what the renderer samples.
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/atmo_consumer.h"

namespace atmo {
atmo_status set_error(atmo_status st, const char *msg);
}

namespace {

atmo_status fail(atmo_status st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return atmo::set_error(st, buf);
}

// Nearest IEEE binary16 value of x (round to nearest even; gradual underflow
// below 2^-14; +-inf from 65520 on, the binary16 overflow threshold), as a float.
float half_round(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    const uint32_t sign = b & 0x80000000u, a = b & 0x7FFFFFFFu;
    uint32_t r;
    if (a >= 0x7F800000u) {
        r = b;                                     // inf, NaN
    } else if (a >= 0x477FF000u) {                 // |x| >= 65520
        r = sign | 0x7F800000u;
    } else if (a >= 0x38800000u) {                 // |x| >= 2^-14: 10 fraction bits kept
        r = sign | ((a + 0x0FFFu + ((a >> 13) & 1u)) & 0xFFFFE000u);
    } else {                                       // subnormal binary16: multiples of 2^-24
        const float q = std::fabs(x) * 16777216.0f;      // exact scaling by 2^24
        const float n = std::nearbyint(q);               // round to nearest even (default mode)
        float v = n * 5.9604644775390625e-08f;           // n 2^-24, exact
        memcpy(&r, &v, 4);
        r |= sign;
    }
    float out;
    memcpy(&out, &r, 4);
    return out;
}

struct Fmt {
    const char *name;
    int channels;
    int half;
};
const Fmt kFormats[] = {{"RGBA16F", 4, 1}, {"RGB16F", 3, 1}, {"RG16F", 2, 1}, {"R16F", 1, 1},
                        {"RGBA32F", 4, 0}, {"RGB32F", 3, 0}, {"RG32F", 2, 0}, {"R32F", 1, 0}};

bool read_file(const char *path, std::string &out) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    char buf[4096];
    size_t n;
    out.clear();
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
    fclose(f);
    return true;
}

// The first texture2D / texture3D element of the document (comments and
// processing instructions skipped) and its attributes.
bool find_element(const std::string &doc, std::string &tag, std::vector<std::pair<std::string, std::string>> &attrs,
                  std::string &err) {
    size_t i = 0;
    while ((i = doc.find('<', i)) != std::string::npos) {
        if (doc.compare(i, 4, "<!--") == 0) {
            const size_t e = doc.find("-->", i + 4);
            if (e == std::string::npos) { err = "unterminated comment"; return false; }
            i = e + 3;
            continue;
        }
        if (doc.compare(i, 2, "<?") == 0) {
            const size_t e = doc.find("?>", i + 2);
            if (e == std::string::npos) { err = "unterminated processing instruction"; return false; }
            i = e + 2;
            continue;
        }
        size_t j = i + 1;
        while (j < doc.size() && (isalnum((unsigned char)doc[j]) || doc[j] == '_')) ++j;
        tag = doc.substr(i + 1, j - i - 1);
        const size_t end = doc.find('>', j);
        if (end == std::string::npos) { err = "unterminated element"; return false; }
        if (tag != "texture2D" && tag != "texture3D") {
            i = end + 1;
            continue;
        }
        // attributes: name="value" (single or double quotes)
        size_t k = j;
        attrs.clear();
        while (k < end) {
            while (k < end && isspace((unsigned char)doc[k])) ++k;
            if (k >= end || doc[k] == '/') break;
            size_t n0 = k;
            while (k < end && doc[k] != '=' && !isspace((unsigned char)doc[k])) ++k;
            const std::string key = doc.substr(n0, k - n0);
            while (k < end && isspace((unsigned char)doc[k])) ++k;
            if (k >= end || doc[k] != '=') { err = "attribute " + key + " has no value"; return false; }
            ++k;
            while (k < end && isspace((unsigned char)doc[k])) ++k;
            if (k >= end || (doc[k] != '"' && doc[k] != '\'')) { err = "attribute " + key + " is not quoted"; return false; }
            const char q = doc[k++];
            const size_t v0 = k;
            while (k < end && doc[k] != q) ++k;
            if (k >= end) { err = "attribute " + key + " is not terminated"; return false; }
            attrs.push_back({key, doc.substr(v0, k - v0)});
            ++k;
        }
        return true;
    }
    err = "no texture2D or texture3D element";
    return false;
}

void copy_str(char *dst, size_t cap, const std::string &s) {
    const size_t n = std::min(cap - 1, s.size());
    memcpy(dst, s.data(), n);
    dst[n] = 0;
}

atmo_status parse(const char *xml_path, atmo_texture_desc &d, std::string &raw_path) {
    memset(&d, 0, sizeof d);
    std::string doc, err, tag;
    if (!read_file(xml_path, doc)) return fail(ATMO_ERR_IO, "cannot open %s: %s", xml_path, strerror(errno));
    std::vector<std::pair<std::string, std::string>> attrs;
    if (!find_element(doc, tag, attrs, err)) return fail(ATMO_ERR_IO, "%s: %s", xml_path, err.c_str());
    d.kind = tag == "texture3D" ? 3 : 2;
    std::string fmt, source;
    int have_wrap[3] = {0, 0, 0};
    for (auto &a : attrs) {
        const std::string &k = a.first, &v = a.second;
        if (k == "name") copy_str(d.name, sizeof d.name, v);
        else if (k == "source") source = v;
        else if (k == "internalformat") fmt = v;
        else if (k == "min") d.min_linear = v == "LINEAR";
        else if (k == "mag") d.mag_linear = v == "LINEAR";
        else if (k == "wraps" || k == "wrapt" || k == "wrapr") {
            const int axis = k == "wraps" ? 0 : (k == "wrapt" ? 1 : 2);
            have_wrap[axis] = 1;
            d.clamp_to_edge[axis] = v == "CLAMP_TO_EDGE";
        }
    }
    if (source.empty()) return fail(ATMO_ERR_IO, "%s: %s has no source", xml_path, tag.c_str());
    d.format_channels = 0;
    for (const Fmt &f : kFormats)
        if (fmt == f.name) {
            d.format_channels = f.channels;
            d.half = f.half;
        }
    if (!d.format_channels)
        return fail(ATMO_ERR_UNSUPPORTED, "%s: internalformat \"%s\" is not a float format this loader knows",
                    xml_path, fmt.c_str());
    for (int a = 0; a < d.kind; ++a)
        if (!have_wrap[a])
            return fail(ATMO_ERR_UNSUPPORTED, "%s: no %s attribute (the wrap mode must be explicit)", xml_path,
                        a == 0 ? "wraps" : (a == 1 ? "wrapt" : "wrapr"));
    copy_str(d.format, sizeof d.format, fmt);
    // the source, relative to the descriptor's directory
    if (source[0] == '/') {
        raw_path = source;
    } else {
        const std::string x(xml_path);
        const size_t slash = x.rfind('/');
        raw_path = (slash == std::string::npos ? std::string() : x.substr(0, slash + 1)) + source;
    }
    copy_str(d.source, sizeof d.source, raw_path);
    int32_t h[5];
    atmo_status st = atmo_read_raw_header(raw_path.c_str(), h);
    if (st != ATMO_OK) return st;
    d.width = h[1];
    d.depth = h[3];
    d.height = h[3] > 0 ? h[2] / h[3] : h[2];
    d.components = h[4];
    if (d.kind == 3 && d.depth <= 0)
        return fail(ATMO_ERR_IO, "%s: texture3D but %s has depth 0 (a 2D table)", xml_path, raw_path.c_str());
    if (d.kind == 2 && d.depth != 0)
        return fail(ATMO_ERR_IO, "%s: texture2D but %s has depth %d (a 3D table)", xml_path, raw_path.c_str(), d.depth);
    if (d.components > d.format_channels)
        return fail(ATMO_ERR_UNSUPPORTED, "%s: %d components do not fit %s", xml_path, d.components, d.format);
    return ATMO_OK;
}

}  // namespace

extern "C" {

atmo_status atmo_texture_load_desc(const char *xml_path, atmo_texture_desc *out) {
    if (!xml_path || !out) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    std::string raw;
    return parse(xml_path, *out, raw);
}

atmo_status atmo_texture_load(const char *xml_path, float *data, size_t capacity_floats, atmo_texture_desc *out) {
    if (!xml_path) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    atmo_texture_desc d;
    std::string raw;
    atmo_status st = parse(xml_path, d, raw);
    if (st != ATMO_OK) return st;
    if (out) *out = d;
    const size_t n = (size_t)d.width * d.height * (d.depth > 0 ? d.depth : 1) * d.components;
    if (!data || capacity_floats < n)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "%s: buffer too small (%zu < %zu floats)", xml_path,
                    capacity_floats, n);
    int32_t h[5];
    if ((st = atmo_read_raw(raw.c_str(), data, capacity_floats, h)) != ATMO_OK) return st;
    if (d.half)
        for (size_t i = 0; i < n; ++i) data[i] = half_round(data[i]);
    return ATMO_OK;
}

}  // extern "C"
