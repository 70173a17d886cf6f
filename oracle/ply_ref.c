/* ply_ref.c -- test infrastructure only: the reference's PLY reader and
 * writer (filePLY.cpp:116-333) restated in C on top of the reference's own
 * rply library (third_party/rply/rply/rply.c, compiled from the reference
 * tree by oracle/Makefile, never copied).  The tests compare the library's
 * own PLY code, which does not use rply, against it.
 *
 *   plyref_read  : readPLY (filePLY.cpp:199-257) with the vertex, normal and
 *                  face callbacks (:116-195) and ear clipping (:12-107)
 *   plyref_write : writePLY (:259-333)
 */
#include <rply.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float    *pts, *nrm;
    long      vnum, vidx, nnum, nidx;
    unsigned *poly;
    long      plen, pcap;
    int      *tri;
    long      ntri, tricap;
    long      fnum, fidx;
} state_t;

/* filePLY.cpp:12-41 on the x/y of a triangle */
static int inside_xy(const float tri[6], double x, double y) {
    int inside = 0;
    for (int i = 0; i < 3; ++i) {
        int    j   = (i + 1) % 3;
        double vx0 = tri[2 * i], vy0 = tri[2 * i + 1], vx1 = tri[2 * j], vy1 = tri[2 * j + 1];
        if (((vy0 <= y) && (vy1 > y)) || ((vy1 <= y) && (vy0 > y))) {
            double cross = (vx1 - vx0) * (y - vy0) / (vy1 - vy0) + vx0;
            if (cross < x)
                inside = !inside;
        }
    }
    return inside;
}

static void push_tri(state_t *s, unsigned a, unsigned b, unsigned c) {
    if (s->ntri == s->tricap) {
        s->tricap = s->tricap ? 2 * s->tricap : 64;
        s->tri    = (int *)realloc(s->tri, sizeof(int) * 3 * (size_t)s->tricap);
    }
    s->tri[3 * s->ntri]     = (int)a;
    s->tri[3 * s->ntri + 1] = (int)b;
    s->tri[3 * s->ntri + 2] = (int)c;
    s->ntri++;
}

/* AddTrianglesByEarClipping, filePLY.cpp:45-107 (Eigen float vectors) */
static int ear_clip(state_t *s, unsigned *idx, int n) {
    const float *P = s->pts;
    float        fn[3] = {0, 0, 0};
    if (n > 3) {
        for (int i = 0; i < n; i++) {
            const float *a = P + 3 * idx[i % n], *b = P + 3 * idx[(i + 1) % n],
                        *c = P + 3 * idx[(i + 2) % n];
            float v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            float v2[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
            fn[0] += v1[1] * v2[2] - v1[2] * v2[1];
            fn[1] += v1[2] * v2[0] - v1[0] * v2[2];
            fn[2] += v1[0] * v2[1] - v1[1] * v2[0];
        }
        float l  = sqrtf(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
        float sc = (float)(1.0 / l);
        for (int k = 0; k < 3; k++)
            fn[k] *= sc;
    }
    int found = 1;
    while (n > 3) {
        if (!found)
            return 0;
        found = 0;
        for (int i = 1; i < n - 1; i++) {
            const float *a = P + 3 * idx[i - 1], *b = P + 3 * idx[i], *c = P + 3 * idx[i + 1];
            float v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            float v2[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
            float x[3]  = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                           v1[0] * v2[1] - v1[1] * v2[0]};
            if (!(fn[0] * x[0] + fn[1] * x[1] + fn[2] * x[2] > 0.0f))
                continue;
            float tri[6] = {a[0], a[1], b[0], b[1], c[0], c[1]};
            int   ear    = 1;
            for (int j = 0; j < n; j++) {
                if (j == i - 1 || j == i || j == i + 1)
                    continue;
                if (inside_xy(tri, P[3 * idx[j]], P[3 * idx[j] + 1])) {
                    ear = 0;
                    break;
                }
            }
            if (ear) {
                found = 1;
                push_tri(s, idx[i - 1], idx[i], idx[i + 1]);
                memmove(idx + i, idx + i + 1, sizeof(unsigned) * (size_t)(n - i - 1));
                n--;
                break;
            }
        }
    }
    push_tri(s, idx[0], idx[1], idx[2]);
    return 1;
}

static int vertex_cb(p_ply_argument a) { /* ReadVertexCallback */
    state_t *s;
    long     index;
    ply_get_argument_user_data(a, (void **)&s, &index);
    if (s->vidx >= s->vnum)
        return 0;
    s->pts[3 * s->vidx + index] = (float)ply_get_argument_value(a);
    if (index == 2)
        s->vidx++;
    return 1;
}

static int normal_cb(p_ply_argument a) { /* ReadNormalCallback */
    state_t *s;
    long     index;
    ply_get_argument_user_data(a, (void **)&s, &index);
    if (s->nidx >= s->nnum)
        return 0;
    s->nrm[3 * s->nidx + index] = (float)ply_get_argument_value(a);
    if (index == 2)
        s->nidx++;
    return 1;
}

static int face_cb(p_ply_argument a) { /* ReadFaceCallBack */
    state_t *s;
    long     dummy, length, index;
    ply_get_argument_user_data(a, (void **)&s, &dummy);
    double value = ply_get_argument_value(a);
    if (s->fidx >= s->fnum)
        return 0;
    ply_get_argument_property(a, NULL, &length, &index);
    if (index == -1) {
        s->plen = 0;
    } else {
        if (s->plen == s->pcap) {
            s->pcap = s->pcap ? 2 * s->pcap : 16;
            s->poly = (unsigned *)realloc(s->poly, sizeof(unsigned) * (size_t)s->pcap);
        }
        s->poly[s->plen++] = (unsigned)(int)value;
    }
    if (s->plen == length) {
        if (s->plen < 3 || !ear_clip(s, s->poly, (int)s->plen))
            return 0;
        s->fidx++;
    }
    return 1;
}

/* 0 on failure; on success *pts (nv x 3), *nrm (nn x 3) and *tri (ntri x 3)
 * are malloc'ed and owned by the caller (plyref_free) */
int plyref_read(const char *path, long *nv, long *nn, long *ntri, float **pts, float **nrm,
                int **tri) {
    state_t s;
    memset(&s, 0, sizeof s);
    p_ply ply = ply_open(path, NULL, 0, NULL);
    if (!ply)
        return 0;
    if (!ply_read_header(ply)) {
        ply_close(ply);
        return 0;
    }
    s.vnum = ply_set_read_cb(ply, "vertex", "x", vertex_cb, &s, 0);
    ply_set_read_cb(ply, "vertex", "y", vertex_cb, &s, 1);
    ply_set_read_cb(ply, "vertex", "z", vertex_cb, &s, 2);
    s.nnum = ply_set_read_cb(ply, "vertex", "nx", normal_cb, &s, 0);
    ply_set_read_cb(ply, "vertex", "ny", normal_cb, &s, 1);
    ply_set_read_cb(ply, "vertex", "nz", normal_cb, &s, 2);
    if (s.vnum <= 0) {
        ply_close(ply);
        return 0;
    }
    s.fnum = ply_set_read_cb(ply, "face", "vertex_indices", face_cb, &s, 0);
    if (s.fnum == 0)
        s.fnum = ply_set_read_cb(ply, "face", "vertex_index", face_cb, &s, 0);
    s.pts = (float *)calloc((size_t)s.vnum * 3, sizeof(float));
    s.nrm = (float *)calloc((size_t)(s.nnum ? s.nnum : 1) * 3, sizeof(float));
    int ok = ply_read(ply);
    ply_close(ply);
    free(s.poly);
    if (!ok) {
        free(s.pts);
        free(s.nrm);
        free(s.tri);
        return 0;
    }
    *nv   = s.vnum;
    *nn   = s.nnum;
    *ntri = s.ntri;
    *pts  = s.pts;
    *nrm  = s.nrm;
    *tri  = s.tri;
    return 1;
}

void plyref_free(void *p) { free(p); }

/* writePLY, filePLY.cpp:259-333; 0 on failure */
int plyref_write(const char *path, const float *pts, const float *nrm, long n, const int *tri,
                 long ntri, int ascii) {
    if (n == 0)
        return 0;
    p_ply ply = ply_create(path, ascii ? PLY_ASCII : PLY_LITTLE_ENDIAN, NULL, 0, NULL);
    if (!ply)
        return 0;
    ply_add_comment(ply, "Created by Open3D");
    ply_add_element(ply, "vertex", n);
    ply_add_property(ply, "x", PLY_FLOAT, PLY_FLOAT, PLY_FLOAT);
    ply_add_property(ply, "y", PLY_FLOAT, PLY_FLOAT, PLY_FLOAT);
    ply_add_property(ply, "z", PLY_FLOAT, PLY_FLOAT, PLY_FLOAT);
    if (nrm) {
        ply_add_property(ply, "nx", PLY_FLOAT, PLY_FLOAT, PLY_FLOAT);
        ply_add_property(ply, "ny", PLY_FLOAT, PLY_FLOAT, PLY_FLOAT);
        ply_add_property(ply, "nz", PLY_FLOAT, PLY_FLOAT, PLY_FLOAT);
    }
    ply_add_element(ply, "face", ntri);
    ply_add_property(ply, "vertex_indices", PLY_LIST, PLY_UCHAR, PLY_UINT);
    if (!ply_write_header(ply)) {
        ply_close(ply);
        return 0;
    }
    for (long i = 0; i < n; i++) {
        ply_write(ply, pts[3 * i]);
        ply_write(ply, pts[3 * i + 1]);
        ply_write(ply, pts[3 * i + 2]);
        if (nrm) {
            ply_write(ply, nrm[3 * i]);
            ply_write(ply, nrm[3 * i + 1]);
            ply_write(ply, nrm[3 * i + 2]);
        }
    }
    for (long i = 0; i < ntri; i++) {
        ply_write(ply, 3);
        ply_write(ply, tri[3 * i]);
        ply_write(ply, tri[3 * i + 1]);
        ply_write(ply, tri[3 * i + 2]);
    }
    ply_close(ply);
    return 1;
}
