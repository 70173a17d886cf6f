// PLY point clouds in and out: the reader and writer of filePLY.cpp (readPLY
// :199-257, writePLY :259-333), written here without rply.
//
// Reading follows rply's rules as readPLY uses them: the header grammar
// (magic "ply" whose line ending decides whether one or two characters
// follow "end_header" in binary files, "format ascii|binary_little_endian|
// binary_big_endian 1.0", comment / obj_info lines, elements and scalar or
// list properties of the types int8..float64 and their C aliases), ASCII
// values as whitespace-separated tokens converted by strtol (integers, with
// the type's range) and strtod (floats, |v| <= FLT_MAX for float32), and
// every value handed over as a double.  readPLY keeps vertex x/y/z, vertex
// nx/ny/nz (when nx is present) and the polygons of face "vertex_indices"
// (else "vertex_index"), splitting polygons into triangles by ear clipping
// (filePLY.cpp:45-107).
#include "ppf_dev.h"
#include "../../include/ppf_c.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ppfd {
namespace {

enum PlyType { T_I8, T_U8, T_I16, T_U16, T_I32, T_U32, T_F32, T_F64, T_LIST, T_BAD };

PlyType ply_type(const std::string &s) {
    static const char *names[] = {"int8", "uint8", "int16", "uint16", "int32", "uint32",
                                  "float32", "float64", "char", "uchar", "short", "ushort",
                                  "int", "uint", "float", "double", "list"};
    static const PlyType types[] = {T_I8,  T_U8,  T_I16, T_U16, T_I32, T_U32, T_F32, T_F64, T_I8,
                                    T_U8,  T_I16, T_U16, T_I32, T_U32, T_F32, T_F64, T_LIST};
    for (int i = 0; i < 17; i++)
        if (s == names[i])
            return types[i];
    return T_BAD;
}
int ply_type_size(PlyType t) {
    static const int s[] = {1, 1, 2, 2, 4, 4, 4, 8};
    return s[t];
}

struct PlyProperty {
    std::string name;
    PlyType     type = T_BAD, lengthType = T_BAD, valueType = T_BAD;
};
struct PlyElement {
    std::string              name;
    long                     count = 0;
    std::vector<PlyProperty> props;
};

// whole-file reader with a cursor; the header is read as tokens, the body
// as tokens (ASCII) or raw bytes (binary)
struct PlyFile {
    std::vector<char> buf;
    size_t            pos = 0;
    int               format = -1;  // 0 ascii, 1 little endian, 2 big endian
    std::vector<PlyElement> elements;

    static bool blank(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t'; }
    // next whitespace-separated token (empty at the end of the file)
    std::string token() {
        while (pos < buf.size() && blank(buf[pos]))
            pos++;
        size_t b = pos;
        while (pos < buf.size() && !blank(buf[pos]))
            pos++;
        std::string t(buf.data() + b, pos - b);
        if (pos < buf.size())
            pos++;  // the delimiter after the token
        return t;
    }
    // the next token of the data, NUL-terminated in place (the buffer
    // carries one spare NUL at its end); empty at the end of the file
    const char *data_token() {
        const size_t size = buf.size() - 1;
        while (pos < size && blank(buf[pos]))
            pos++;
        const size_t b = pos;
        while (pos < size && !blank(buf[pos]))
            pos++;
        if (pos < size)
            buf[pos++] = '\0';
        return buf.data() + b;
    }
    // from the cursor (past the delimiter of the token just read) through
    // the next newline, as rply reads a comment line
    void skip_line() {
        while (pos < buf.size() && buf[pos] != '\n')
            pos++;
        if (pos < buf.size())
            pos++;
    }
};

[[noreturn]] void ply_fail(const std::string &what) {
    throw std::runtime_error("Read PLY failed: " + what);
}

void read_header(PlyFile &f) {
    const std::vector<char> &b = f.buf;
    if (b.size() < 5 || b[0] != 'p' || b[1] != 'l' || b[2] != 'y' ||
        !(b[3] == ' ' || b[3] == '\n' || b[3] == '\r' || b[3] == '\t' || b[3] == '\v' ||
          b[3] == '\f'))
        ply_fail("unable to parse header.");
    const bool crlf = b.size() > 4 && b[3] == '\r' && b[4] == '\n';
    f.pos           = 3;
    std::string t   = f.token();
    if (t != "format")
        ply_fail("unable to parse header.");
    t = f.token();
    if (t == "ascii")
        f.format = 0;
    else if (t == "binary_little_endian")
        f.format = 1;
    else if (t == "binary_big_endian")
        f.format = 2;
    else
        ply_fail("unable to parse header.");
    if (f.token() != "1.0")
        ply_fail("unable to parse header.");
    t = f.token();
    for (;;) {
        if (t == "end_header")
            break;
        if (t == "comment" || t == "obj_info") {
            f.skip_line();
            t = f.token();
        } else if (t == "element") {
            PlyElement e;
            e.name            = f.token();
            const std::string n = f.token();
            char             *end = nullptr;
            e.count           = std::strtol(n.c_str(), &end, 10);
            // a negative count would size the reference's arrays from a
            // negative long (filePLY.cpp:211-216); refused here
            if (n.empty() || end == n.c_str() || e.count < 0)
                ply_fail("unable to parse header.");
            t = f.token();
            while (t == "property" || t == "comment" || t == "obj_info") {
                if (t != "property") {
                    f.skip_line();
                    t = f.token();
                    continue;
                }
                PlyProperty p;
                p.type = ply_type(f.token());
                if (p.type == T_BAD)
                    ply_fail("unable to parse header.");
                if (p.type == T_LIST) {
                    p.lengthType = ply_type(f.token());
                    p.valueType  = ply_type(f.token());
                    if (p.lengthType == T_BAD || p.lengthType == T_LIST || p.valueType == T_BAD ||
                        p.valueType == T_LIST)
                        ply_fail("unable to parse header.");
                }
                p.name = f.token();
                e.props.push_back(p);
                t = f.token();
            }
            f.elements.push_back(e);
        } else {
            ply_fail("unable to parse header.");
        }
    }
    // binary data start right after "end_header" and its line ending
    if (crlf && f.format != 0)
        f.pos++;
}

// one value of type t, as a double
double read_value(PlyFile &f, PlyType t) {
    if (f.format == 0) {
        const char *tok = f.data_token();
        if (!*tok)
            ply_fail("unable to read file.");
        char *end = nullptr;
        if (t == T_F32 || t == T_F64) {
            const double v = std::strtod(tok, &end);
            const double m = t == T_F32 ? FLT_MAX : DBL_MAX;
            if (*end || v < -m || v > m)
                ply_fail("unable to read file.");
            return v;
        }
        const double v = (double)std::strtol(tok, &end, 10);
        static const double lo[] = {-128, 0, -32768, 0, -2147483648.0, 0};
        static const double hi[] = {127, 255, 32767, 65535, 2147483647.0, 4294967295.0};
        if (*end || v < lo[t] || v > hi[t])
            ply_fail("unable to read file.");
        return v;
    }
    const int n = ply_type_size(t);
    if (f.pos + n > f.buf.size() - 1)
        ply_fail("unable to read file.");
    unsigned char raw[8];
    std::memcpy(raw, f.buf.data() + f.pos, n);
    f.pos += n;
    if (f.format == 2)
        for (int i = 0; i < n / 2; i++)
            std::swap(raw[i], raw[n - 1 - i]);
    switch (t) {
        case T_I8: { int8_t v; std::memcpy(&v, raw, 1); return v; }
        case T_U8: { uint8_t v; std::memcpy(&v, raw, 1); return v; }
        case T_I16: { int16_t v; std::memcpy(&v, raw, 2); return v; }
        case T_U16: { uint16_t v; std::memcpy(&v, raw, 2); return v; }
        case T_I32: { int32_t v; std::memcpy(&v, raw, 4); return v; }
        case T_U32: { uint32_t v; std::memcpy(&v, raw, 4); return v; }
        case T_F32: { float v; std::memcpy(&v, raw, 4); return v; }
        default: { double v; std::memcpy(&v, raw, 8); return v; }
    }
}

// point-in-triangle test on x/y only (filePLY.cpp:12-41)
bool inside_xy(const float *tri, double x, double y) {
    bool inside = false;
    for (int i = 0; i < 3; ++i) {
        const int    j   = (i + 1) % 3;
        const double vx0 = tri[2 * i], vy0 = tri[2 * i + 1], vx1 = tri[2 * j], vy1 = tri[2 * j + 1];
        if (((vy0 <= y) && (vy1 > y)) || ((vy1 <= y) && (vy0 > y))) {
            const double cross = (vx1 - vx0) * (y - vy0) / (vy1 - vy0) + vx0;
            if (cross < x)
                inside = !inside;
        }
    }
    return inside;
}

void cross3(const float *a, const float *b, float *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// ear clipping of one polygon into triangles (filePLY.cpp:45-107); false
// if no ear is left
bool triangulate(const std::vector<float> &pts, std::vector<unsigned> idx, std::vector<int> &faces) {
    int  n = (int)idx.size();
    auto P = [&](int i) { return &pts[3 * (size_t)idx[i]]; };
    float nrm[3] = {0.0f, 0.0f, 0.0f};
    if (n > 3) {
        for (int i = 0; i < n; i++) {
            const float *a = P(i % n), *b = P((i + 1) % n), *c = P((i + 2) % n);
            float v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            float v2[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]}, x[3];
            cross3(v1, v2, x);
            for (int k = 0; k < 3; k++)
                nrm[k] = nrm[k] + x[k];
        }
        const float l = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        const float s = (float)(1.0 / l);
        for (int k = 0; k < 3; k++)
            nrm[k] = nrm[k] * s;
    }
    bool found = true;
    while (n > 3) {
        if (!found)
            return false;
        found = false;
        for (int i = 1; i < n - 1; i++) {
            const float *a = P(i - 1), *b = P(i), *c = P(i + 1);
            float v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            float v2[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]}, x[3];
            cross3(v1, v2, x);
            const bool convex = nrm[0] * x[0] + nrm[1] * x[1] + nrm[2] * x[2] > 0.0f;
            if (!convex)
                continue;
            const float tri[6] = {a[0], a[1], b[0], b[1], c[0], c[1]};
            bool        ear    = true;
            for (int j = 0; j < n; j++) {
                if (j == i - 1 || j == i || j == i + 1)
                    continue;
                if (inside_xy(tri, P(j)[0], P(j)[1])) {
                    ear = false;
                    break;
                }
            }
            if (ear) {
                found = true;
                faces.push_back((int)idx[i - 1]);
                faces.push_back((int)idx[i]);
                faces.push_back((int)idx[i + 1]);
                idx.erase(idx.begin() + i);
                n = (int)idx.size();
                break;
            }
        }
    }
    faces.push_back((int)idx[0]);
    faces.push_back((int)idx[1]);
    faces.push_back((int)idx[2]);
    return true;
}

}  // namespace

void ply_read(const char *path, PlyCloud &out) {
    PlyFile f;
    {
        std::unique_ptr<FILE, int (*)(FILE *)> fp(std::fopen(path, "rb"), &std::fclose);
        if (!fp)
            ply_fail(std::string("unable to open file: ") + path);
        std::fseek(fp.get(), 0, SEEK_END);
        const long size = std::ftell(fp.get());
        std::fseek(fp.get(), 0, SEEK_SET);
        f.buf.resize(size > 0 ? (size_t)size : 0);
        if (size > 0 && std::fread(f.buf.data(), 1, (size_t)size, fp.get()) != (size_t)size)
            ply_fail(std::string("unable to read file: ") + path);
        f.buf.push_back('\0');  // terminates the last ASCII token
    }
    read_header(f);
    // the properties readPLY asks for (filePLY.cpp:207-225)
    const PlyElement *vertex = nullptr;
    for (const PlyElement &e : f.elements)
        if (e.name == "vertex" && !vertex)
            vertex = &e;
    auto has = [](const PlyElement *e, const char *p) {
        if (!e)
            return false;
        for (const PlyProperty &q : e->props)
            if (q.name == p)
                return true;
        return false;
    };
    const long nv = has(vertex, "x") ? vertex->count : 0;
    if (nv <= 0)
        ply_fail("number of vertex <= 0.");
    const long nn = has(vertex, "nx") ? vertex->count : 0;
    out.points.assign((size_t)nv * 3, 0.0f);
    out.normals.assign((size_t)nn * 3, 0.0f);
    out.faces.clear();
    for (const PlyElement &e : f.elements) {
        const bool isVertex = &e == vertex;
        const bool isFace   = e.name == "face";
        const char *faceProp = nullptr;
        if (isFace)
            faceProp = has(&e, "vertex_indices") ? "vertex_indices"
                       : has(&e, "vertex_index") ? "vertex_index"
                                                 : nullptr;
        // where each scalar property of the vertex element goes: 0-2 the
        // point, 3-5 the normal, -1 nowhere
        std::vector<int> slot(e.props.size(), -1);
        bool             lists = false;
        for (size_t q = 0; q < e.props.size(); q++) {
            const std::string &nm = e.props[q].name;
            lists                 = lists || e.props[q].type == T_LIST;
            if (!isVertex || e.props[q].type == T_LIST)
                continue;
            if (nm == "x") slot[q] = 0;
            else if (nm == "y") slot[q] = 1;
            else if (nm == "z") slot[q] = 2;
            else if (nn && nm == "nx") slot[q] = 3;
            else if (nn && nm == "ny") slot[q] = 4;
            else if (nn && nm == "nz") slot[q] = 5;
        }
        if (f.format != 0 && !lists) {
            // binary rows of scalars: fixed stride, values converted in place
            size_t stride = 0;
            std::vector<size_t> off(e.props.size());
            for (size_t q = 0; q < e.props.size(); q++) {
                off[q] = stride;
                stride += (size_t)ply_type_size(e.props[q].type);
            }
            // the rows must fit in what is left of the file (compared by
            // division: stride * count may not fit a size_t)
            const size_t left = f.buf.size() - 1 > f.pos ? f.buf.size() - 1 - f.pos : 0;
            if (stride > 0 && (size_t)e.count > left / stride)
                ply_fail("unable to read file.");
            if (isVertex) {
                for (long i = 0; i < e.count; i++) {
                    const char *row = f.buf.data() + f.pos + stride * (size_t)i;
                    for (size_t q = 0; q < e.props.size(); q++) {
                        if (slot[q] < 0)
                            continue;
                        const int n = ply_type_size(e.props[q].type);
                        unsigned char raw[8];
                        std::memcpy(raw, row + off[q], n);
                        if (f.format == 2)
                            for (int b = 0; b < n / 2; b++)
                                std::swap(raw[b], raw[n - 1 - b]);
                        double v;
                        switch (e.props[q].type) {
                            case T_I8: { int8_t w; std::memcpy(&w, raw, 1); v = w; break; }
                            case T_U8: { uint8_t w; std::memcpy(&w, raw, 1); v = w; break; }
                            case T_I16: { int16_t w; std::memcpy(&w, raw, 2); v = w; break; }
                            case T_U16: { uint16_t w; std::memcpy(&w, raw, 2); v = w; break; }
                            case T_I32: { int32_t w; std::memcpy(&w, raw, 4); v = w; break; }
                            case T_U32: { uint32_t w; std::memcpy(&w, raw, 4); v = w; break; }
                            case T_F32: { float w; std::memcpy(&w, raw, 4); v = w; break; }
                            default: { double w; std::memcpy(&w, raw, 8); v = w; break; }
                        }
                        (slot[q] < 3 ? out.points : out.normals)[3 * (size_t)i + slot[q] % 3] =
                            (float)v;
                    }
                }
            }
            f.pos += stride * (size_t)e.count;
            continue;
        }
        std::vector<unsigned> poly;
        for (long i = 0; i < e.count; i++) {
            for (size_t q = 0; q < e.props.size(); q++) {
                const PlyProperty &p = e.props[q];
                if (p.type == T_LIST) {
                    const double len  = read_value(f, p.lengthType);
                    const bool   keep = isFace && faceProp && p.name == faceProp;
                    poly.clear();
                    for (long j = 0; j < (long)len; j++) {
                        const double v = read_value(f, p.valueType);
                        if (keep)
                            poly.push_back((unsigned)(int)v);
                    }
                    if (keep) {
                        if (poly.size() < 3)
                            ply_fail("a polygon has fewer than 3 vertices.");
                        // a triangle is kept as it is; larger polygons read
                        // their vertices to be split
                        if (poly.size() > 3)
                            for (unsigned q : poly)
                                if (q >= (unsigned)nv)
                                    ply_fail("a face refers to a missing vertex.");
                        if (!triangulate(out.points, poly, out.faces))
                            ply_fail("A polygon in the mesh could not be decomposed into triangles.");
                    }
                    continue;
                }
                const double v = read_value(f, p.type);
                if (slot[q] >= 0)
                    (slot[q] < 3 ? out.points : out.normals)[3 * (size_t)i + slot[q] % 3] =
                        (float)v;
            }
        }
    }
}

// writePLY: the header rply writes for it, then float32 vertices (and
// normals) and uchar-counted uint triangles, little endian or ASCII
// ("%g" of the float, one instance per line)
void ply_write(const char *path, const float *pts, const float *nrm, size_t n, const int *faces,
               size_t nFaces, bool ascii) {
    if (n == 0)
        throw std::runtime_error("Write PLY failed: mesh has 0 vertices.");
    std::unique_ptr<FILE, int (*)(FILE *)> fp(std::fopen(path, "wb"), &std::fclose);
    if (!fp)
        throw std::runtime_error(std::string("Write PLY failed: unable to open file: ") + path);
    FILE *o = fp.get();
    std::fprintf(o, "ply\nformat %s 1.0\ncomment Created by Open3D\nelement vertex %ld\n",
                 ascii ? "ascii" : "binary_little_endian", (long)n);
    std::fprintf(o, "property float x\nproperty float y\nproperty float z\n");
    if (nrm)
        std::fprintf(o, "property float nx\nproperty float ny\nproperty float nz\n");
    std::fprintf(o, "element face %ld\nproperty list uchar uint vertex_indices\nend_header\n",
                 (long)nFaces);
    const int per = nrm ? 6 : 3;
    if (ascii) {
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < per; k++) {
                const float v = k < 3 ? pts[3 * i + k] : nrm[3 * i + k - 3];
                std::fprintf(o, "%g%c", (double)v, k + 1 < per ? ' ' : '\n');
            }
        for (size_t i = 0; i < nFaces; i++)
            std::fprintf(o, "3 %d %d %d\n", (int)(uint32_t)faces[3 * i],
                         (int)(uint32_t)faces[3 * i + 1], (int)(uint32_t)faces[3 * i + 2]);
    } else {
        std::vector<float> row(per);
        for (size_t i = 0; i < n; i++) {
            for (int k = 0; k < per; k++)
                row[k] = k < 3 ? pts[3 * i + k] : nrm[3 * i + k - 3];
            std::fwrite(row.data(), sizeof(float), per, o);
        }
        for (size_t i = 0; i < nFaces; i++) {
            const unsigned char c = 3;
            const uint32_t      t[3] = {(uint32_t)faces[3 * i], (uint32_t)faces[3 * i + 1],
                                        (uint32_t)faces[3 * i + 2]};
            std::fwrite(&c, 1, 1, o);
            std::fwrite(t, sizeof(uint32_t), 3, o);
        }
    }
    if (std::ferror(o))
        throw std::runtime_error(std::string("Write PLY failed: unable to write file: ") + path);
}

}  // namespace ppfd
