// gr_ply.cpp -- 3DGS-style binary PLY scenes <-> the packed host model
// (include/gr_capi.h: gr_ply_probe / gr_ply_read / gr_ply_write).
//
// Reading follows core/ply_parser.cpp:
//   header   parseHeader (:15-100): one token-separated line at a time;
//            "ply" lines are skipped, "format" must name binary_little_endian,
//            the only element is "vertex" with its count, every property is
//            "float" and unique, other keywords (comment, obj_info) are
//            ignored, "end_header" ends it, more than 256 counted lines is an
//            error.
//   data     parsePly (:102-131): vertex-major rows of 4-byte floats in
//            header order.
//   model    parsePly (:241-393): x y z, rot_0..rot_3 with rot_0 = w
//            (glm::quat(w, x, y, z)), scale_0..2, f_dc_0..2, opacity and
//            f_rest_0.. (3 * ((d+1)^2 - 1) of them fix the SH degree d);
//            quaternion normalised as glm::normalize (dot as
//            (w*w + x*x) + (y*y + z*z), 1/len, identity when len <= 0),
//            scale = exp, opacity = sigmoid, rgb = f_dc * SH_C0 + 0.5,
//            rotation packed by encodeVec4((q.xyzw + 1) * 0.5); SH as
//            [N][3 + extra] with the planar f_rest blocks interleaved so that
//            coefficient j, channel c sits at 3j + c.
//   bounds   min/max of the positions, with the maximum starting from
//            -FLT_MAX (the reference starts it from FLT_MIN,
//            core/ply_parser.cpp:352, which clips scenes lying below 0).
// Writing follows savePly (:143-219), including its rotation order: the
// decoded (x, y, z, w) goes out as rot_0..rot_3, so a file written here and
// read back has its quaternion components rotated -- as in the reference.
//
// Differences kept on purpose: read errors (short file, unopened file) are
// reported instead of leaving unread values; a missing end_header is an error.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gr_capi.h"

namespace gr_detail {
void set_last_error(const std::string& msg);
}

namespace {

constexpr float kShC0 = 0.28209479177387814f;

struct PlyError {
    gr_status status;
    std::string msg;
};

struct Header {
    std::vector<std::string> props;
    int64_t count = -1;
    std::streamoff data_offset = 0;
};

Header parse_header(std::ifstream& in) {
    Header h;
    bool little = false;
    constexpr int kMaxLines = 256;
    int iteration = 0;
    std::string line;
    bool ended = false;
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        std::string word;
        if (!(ss >> word)) throw PlyError{GR_ERR_INVALID_ARGUMENT, "Unexpected empty line."};
        if (word == "ply") continue;
        if (word == "format") {
            little = (ss >> word) && word == "binary_little_endian";
        } else if (word == "element") {
            if (!(ss >> word && word == "vertex")) throw PlyError{GR_ERR_INVALID_ARGUMENT, "Unexpected element type."};
            if (!(ss >> word)) throw PlyError{GR_ERR_INVALID_ARGUMENT, "Unexpected element count."};
            try {
                size_t used = 0;
                const long long v = std::stoll(word, &used);
                h.count = v;
            } catch (const std::exception&) {
                throw PlyError{GR_ERR_INVALID_ARGUMENT, "Unexpected element count."};
            }
            if (h.count < 0 || h.count > 0x7fffffffLL)
                throw PlyError{GR_ERR_INVALID_ARGUMENT, "Unexpected element count."};
        } else if (word == "property") {
            if (!(ss >> word && word == "float"))
                throw PlyError{GR_ERR_INVALID_ARGUMENT, "Unexpected property format, expected float."};
            if (!(ss >> word)) throw PlyError{GR_ERR_INVALID_ARGUMENT, "Unexpected property, missing name."};
            if (std::find(h.props.begin(), h.props.end(), word) != h.props.end())
                throw PlyError{GR_ERR_INVALID_ARGUMENT, "Duplicated property \"" + word + "\"."};
            h.props.push_back(word);
        } else if (word == "end_header") {
            if (!little || h.count == -1)
                throw PlyError{GR_ERR_INVALID_ARGUMENT, "Invalid header. Format or vertex count not found."};
            ended = true;
            break;
        }
        ++iteration;
        if (iteration > kMaxLines) throw PlyError{GR_ERR_INVALID_ARGUMENT, "Invalid header. End not found."};
    }
    if (!ended) throw PlyError{GR_ERR_INVALID_ARGUMENT, "Invalid header. End not found."};
    h.data_offset = in.tellg();
    return h;
}

int index_of(const std::vector<std::string>& props, const std::string& name) {
    auto it = std::find(props.begin(), props.end(), name);
    if (it == props.end()) throw PlyError{GR_ERR_INVALID_ARGUMENT, "Missing property \"" + name + "\"."};
    return (int)std::distance(props.begin(), it);
}

struct Layout {
    int off[14];
    std::vector<int> rest;
    int degree = 0;
    int sh_count = 3;
};

Layout layout_of(const Header& h) {
    static const char* kNames[14] = {"x",       "y",       "z",      "rot_0",  "rot_1",  "rot_2",  "rot_3",
                                     "scale_0", "scale_1", "scale_2", "f_dc_0", "f_dc_1", "f_dc_2", "opacity"};
    Layout l;
    for (int k = 0; k < 14; ++k) l.off[k] = index_of(h.props, kNames[k]);
    for (int j = 0;; ++j) {
        auto it = std::find(h.props.begin(), h.props.end(), "f_rest_" + std::to_string(j));
        if (it == h.props.end()) break;
        l.rest.push_back((int)std::distance(h.props.begin(), it));
    }
    const int extra = (int)l.rest.size();
    int expected = 0, degree = 0;
    while (expected < extra) {
        expected += (2 * (degree + 1) + 1) * 3;
        ++degree;
    }
    if (expected != extra)
        throw PlyError{GR_ERR_INVALID_ARGUMENT, "Expected degree " + std::to_string(degree) + ", " +
                                                    std::to_string(expected) + " extra spherical harmonics." +
                                                    " Found " + std::to_string(extra) + "."};
    if (degree > 3)
        throw PlyError{GR_ERR_INVALID_ARGUMENT, "SH degree " + std::to_string(degree) + " is above 3"};
    l.degree = degree;
    l.sh_count = extra + 3;
    return l;
}

std::ifstream open_ply(const char* path) {
    std::ifstream in(path, std::ios_base::binary);
    if (!in.is_open()) throw PlyError{GR_ERR_NOT_FOUND, std::string("cannot open ") + path + ": " + std::strerror(errno)};
    return in;
}

// encodeVec4 (core/ply_parser.cpp:229-239): clamp to [0, 1], truncate x*255
uint32_t encode_vec4(float x, float y, float z, float w) {
    auto q = [](float v) { return (uint32_t)(std::min(std::max(v, 0.0f), 1.0f) * 255.0f); };
    return (q(x) << 24u) | (q(y) << 16u) | (q(z) << 8u) | q(w);
}

// decodeVec4 (core/ply_parser.cpp:133-141)
void decode_vec4(uint32_t e, float out[4]) {
    out[0] = (float)((e >> 24u) & 0xffu) / 255.0f * 2.0f - 1.0f;
    out[1] = (float)((e >> 16u) & 0xffu) / 255.0f * 2.0f - 1.0f;
    out[2] = (float)((e >> 8u) & 0xffu) / 255.0f * 2.0f - 1.0f;
    out[3] = (float)(e & 0xffu) / 255.0f * 2.0f - 1.0f;
}

float sigmoid(float x) { return 1.0f / (1.0f + std::exp(-x)); }

}  // namespace

extern "C" {

gr_status gr_ply_probe(const char* path, int32_t* count, int32_t* sh_degree, int32_t* sh_count) {
    if (!path) {
        gr_detail::set_last_error("gr_ply_probe: path is NULL");
        return GR_ERR_INVALID_ARGUMENT;
    }
    try {
        std::ifstream in = open_ply(path);
        const Header h = parse_header(in);
        const Layout l = layout_of(h);
        if (count) *count = (int32_t)h.count;
        if (sh_degree) *sh_degree = l.degree;
        if (sh_count) *sh_count = l.sh_count;
        return GR_OK;
    } catch (const PlyError& e) {
        gr_detail::set_last_error(e.msg);
        return e.status;
    }
}

gr_status gr_ply_read(const char* path, int32_t capacity, float* position, float* scale_and_rotation, float* color,
                      float* spherical_harmonics, float* rotation_wxyz, float* bounds_min, float* bounds_max,
                      int32_t* count) {
    if (!path || !position || !scale_and_rotation || !color) {
        gr_detail::set_last_error("gr_ply_read: NULL argument");
        return GR_ERR_INVALID_ARGUMENT;
    }
    try {
        std::ifstream in = open_ply(path);
        const Header h = parse_header(in);
        const Layout l = layout_of(h);
        const int64_t n = h.count;
        if (n > capacity)
            throw PlyError{GR_ERR_CAPACITY, "file holds " + std::to_string(n) + " vertices, capacity is " +
                                                std::to_string(capacity)};
        if (l.degree > 0 && !spherical_harmonics)
            throw PlyError{GR_ERR_INVALID_ARGUMENT, "file has SH degree " + std::to_string(l.degree) +
                                                        " but the SH output is NULL"};
        const size_t np = h.props.size();
        const int extra = (int)l.rest.size();
        const int rgb = extra / 3;
        float mn[3], mx[3];
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::numeric_limits<float>::max();
            mx[k] = std::numeric_limits<float>::lowest();
        }
        constexpr int64_t kRows = 16384;
        std::vector<float> buf((size_t)kRows * np);
        for (int64_t base = 0; base < n; base += kRows) {
            const int64_t rows = std::min(kRows, n - base);
            in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(rows * np * sizeof(float)));
            if (in.gcount() != (std::streamsize)(rows * np * sizeof(float)))
                throw PlyError{GR_ERR_INVALID_ARGUMENT, "file ends after " + std::to_string(base) +
                                                            " of " + std::to_string(n) + " vertices"};
            for (int64_t r = 0; r < rows; ++r) {
                const float* row = buf.data() + (size_t)r * np;
                const int64_t i = base + r;
                const float tx = row[l.off[0]], ty = row[l.off[1]], tz = row[l.off[2]];
                // glm::quat(w, x, y, z) from rot_0..rot_3, then glm::normalize
                float qw = row[l.off[3]], qx = row[l.off[4]], qy = row[l.off[5]], qz = row[l.off[6]];
                const float len = std::sqrt((qw * qw + qx * qx) + (qy * qy + qz * qz));
                if (len <= 0.0f) {
                    qw = 1.0f;
                    qx = qy = qz = 0.0f;
                } else {
                    const float inv = 1.0f / len;
                    qw = qw * inv;
                    qx = qx * inv;
                    qy = qy * inv;
                    qz = qz * inv;
                }
                const float sx = std::exp(row[l.off[7]]), sy = std::exp(row[l.off[8]]), sz = std::exp(row[l.off[9]]);
                const float dc0 = row[l.off[10]], dc1 = row[l.off[11]], dc2 = row[l.off[12]];
                const float op = sigmoid(row[l.off[13]]);
                mn[0] = tx < mn[0] ? tx : mn[0];
                mn[1] = ty < mn[1] ? ty : mn[1];
                mn[2] = tz < mn[2] ? tz : mn[2];
                mx[0] = mx[0] < tx ? tx : mx[0];
                mx[1] = mx[1] < ty ? ty : mx[1];
                mx[2] = mx[2] < tz ? tz : mx[2];
                uint32_t rot =
                    encode_vec4((qx + 1.0f) * 0.5f, (qy + 1.0f) * 0.5f, (qz + 1.0f) * 0.5f, (qw + 1.0f) * 0.5f);
                float rotf;
                std::memcpy(&rotf, &rot, 4);
                float* p = position + 4 * i;
                p[0] = tx; p[1] = ty; p[2] = tz; p[3] = op;
                float* s = scale_and_rotation + 4 * i;
                s[0] = sx; s[1] = sy; s[2] = sz; s[3] = rotf;
                float* c = color + 4 * i;
                c[0] = dc0 * kShC0 + 0.5f; c[1] = dc1 * kShC0 + 0.5f; c[2] = dc2 * kShC0 + 0.5f; c[3] = op;
                if (rotation_wxyz) {
                    float* q = rotation_wxyz + 4 * i;
                    q[0] = qw; q[1] = qx; q[2] = qy; q[3] = qz;
                }
                if (l.degree > 0) {
                    float* sh = spherical_harmonics + (size_t)i * l.sh_count;
                    sh[0] = dc0; sh[1] = dc1; sh[2] = dc2;
                    for (int j = 0; j < rgb; ++j) {
                        sh[3 + 3 * j + 0] = row[l.rest[j]];
                        sh[3 + 3 * j + 1] = row[l.rest[rgb + j]];
                        sh[3 + 3 * j + 2] = row[l.rest[2 * rgb + j]];
                    }
                }
            }
        }
        if (bounds_min) std::memcpy(bounds_min, mn, sizeof mn);
        if (bounds_max) std::memcpy(bounds_max, mx, sizeof mx);
        if (count) *count = (int32_t)n;
        return GR_OK;
    } catch (const PlyError& e) {
        gr_detail::set_last_error(e.msg);
        return e.status;
    } catch (const std::bad_alloc&) {
        gr_detail::set_last_error("gr_ply_read: out of host memory");
        return GR_ERR_OUT_OF_MEMORY;
    }
}

gr_status gr_ply_write(const char* path, int32_t count, const float* position, const float* scale_and_rotation,
                       const float* color) {
    if (!path || count <= 0 || !position || !scale_and_rotation || !color) {
        gr_detail::set_last_error("gr_ply_write: bad argument");
        return GR_ERR_INVALID_ARGUMENT;
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        gr_detail::set_last_error(std::string("cannot create ") + path + ": " + std::strerror(errno));
        return GR_ERR_NOT_FOUND;
    }
    std::string hdr = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(count) + "\n";
    for (const char* p : {"x", "y", "z", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0", "scale_1", "scale_2",
                          "f_dc_0", "f_dc_1", "f_dc_2", "opacity"})
        hdr += std::string("property float ") + p + "\n";
    hdr += "end_header\n";
    bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    constexpr int kRows = 16384;
    std::vector<float> buf((size_t)kRows * 14);
    for (int base = 0; ok && base < count; base += kRows) {
        const int rows = std::min(kRows, count - base);
        for (int r = 0; r < rows; ++r) {
            const int i = base + r;
            const float* p = position + 4 * (size_t)i;
            const float* s = scale_and_rotation + 4 * (size_t)i;
            const float* c = color + 4 * (size_t)i;
            float* o = buf.data() + (size_t)r * 14;
            o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
            uint32_t enc;
            std::memcpy(&enc, &s[3], 4);
            decode_vec4(enc, o + 3);
            o[7] = std::log(s[0]); o[8] = std::log(s[1]); o[9] = std::log(s[2]);
            o[10] = (c[0] - 0.5f) / kShC0; o[11] = (c[1] - 0.5f) / kShC0; o[12] = (c[2] - 0.5f) / kShC0;
            const float op = std::min(std::max(p[3], 0.001f), 0.999f);
            o[13] = std::log(op / (1.0f - op));
        }
        ok = std::fwrite(buf.data(), sizeof(float) * 14, (size_t)rows, f) == (size_t)rows;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        gr_detail::set_last_error(std::string("write to ") + path + " failed");
        return GR_ERR_INVALID_ARGUMENT;
    }
    return GR_OK;
}

}  // extern "C"
