/*
 * orc_grid.c -- restatement of the uniform-grid FRNN (gridCount2d.h) and a
 * brute-force reference (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 *   cell coordinates  gridCount2d.h:94-102  x = (int)((p.x - min.x) / cs.x) % nx
 *   cell key          gridCount2d.h:89-92   (y * nx + x) % tot, unsigned
 *   counting sort     gridCount2d.h:137-172 (here stable: original order kept
 *                                            inside a cell; the reference's
 *                                            atomic order is nondeterministic)
 *   query             gridCount2d.h:262-295 3x3 cells, bound test
 *                                            hh.x > nx || hh.x < 0 (sic),
 *                                            r2 = dot(pj - pi, pj - pi) <= rad*rad
 * FP32 arithmetic without contraction (the build uses -ffp-contract=off).
 * The periodic variant (no reference counterpart) wraps cells and uses the
 * minimum-image difference d -= 1 if d > 0.5, d += 1 if d < -0.5 in FP32.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

static void cell_xy(float px, float py, float minx, float miny, float csx, float csy, int nx,
                    int ny, int* cx, int* cy) {
    float qx = px - minx, qy = py - miny;
    int x = (int)(qx / csx), y = (int)(qy / csy);
    x %= nx;
    y %= ny;
    *cx = x;
    *cy = y;
}
static unsigned cell_hash(unsigned x, unsigned y, int nx, int tot) {
    return (y * (unsigned)nx + x) % (unsigned)tot;
}

int orc_grid_sort(const float* xy, int n, float minx, float miny, float csx, float csy, int ncx,
                  int ncy, int* hash, int* perm, int* cell_start) {
    int tot = ncx * ncy;
    int* count = (int*)calloc((size_t)tot + 1, sizeof(int));
    for (int i = 0; i < n; i++) {
        int cx, cy;
        cell_xy(xy[2 * i], xy[2 * i + 1], minx, miny, csx, csy, ncx, ncy, &cx, &cy);
        hash[i] = (int)cell_hash((unsigned)cx, (unsigned)cy, ncx, tot);
        count[hash[i]]++;
    }
    int acc = 0;
    for (int c = 0; c < tot; c++) { cell_start[c] = acc; acc += count[c]; }
    cell_start[tot] = acc;
    int* fill = (int*)malloc(sizeof(int) * (size_t)(tot + 1));
    memcpy(fill, cell_start, sizeof(int) * (size_t)(tot + 1));
    for (int i = 0; i < n; i++) perm[fill[hash[i]]++] = i;
    free(fill);
    free(count);
    return 0;
}

static inline float min_image(float d) {
    if (d > 0.5f) d -= 1.0f;
    else if (d < -0.5f) d += 1.0f;
    return d;
}

int orc_grid_frnn(const float* xy, int n, float minx, float miny, float csx, float csy, int ncx,
                  int ncy, float rad, int periodic, int* counts, int* nbr, int cap) {
    int tot = ncx * ncy;
    int* hash = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int* perm = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int* start = (int*)malloc(sizeof(int) * (size_t)(tot + 1));
    orc_grid_sort(xy, n, minx, miny, csx, csy, ncx, ncy, hash, perm, start);
    float rad2 = rad * rad;
    for (int s = 0; s < n; s++) {
        int i = perm[s];
        float pix = xy[2 * i], piy = xy[2 * i + 1];
        int hx, hy;
        cell_xy(pix, piy, minx, miny, csx, csy, ncx, ncy, &hx, &hy);
        int c = 0;
        for (int ddx = -1; ddx <= 1; ddx++)
            for (int ddy = -1; ddy <= 1; ddy++) {
                int gx = hx + ddx, gy = hy + ddy;
                if (periodic) {
                    gx = (gx + ncx) % ncx;
                    gy = (gy + ncy) % ncy;
                } else {
                    if (gx > ncx || gx < 0) continue;
                    if (gy > ncy || gy < 0) continue;
                }
                unsigned h = cell_hash((unsigned)gx, (unsigned)gy, ncx, tot);
                for (int q = start[h]; q < start[h + 1]; q++) {
                    int j = perm[q];
                    float dx = xy[2 * j] - pix, dy = xy[2 * j + 1] - piy;
                    if (periodic) { dx = min_image(dx); dy = min_image(dy); }
                    float r2 = dx * dx + dy * dy;
                    if (r2 <= rad2 && j != i) {
                        if (c < cap) nbr[(size_t)i * cap + c] = j;
                        c++;
                    }
                }
            }
        counts[i] = c;
    }
    free(hash); free(perm); free(start);
    return 0;
}

int orc_brute_frnn(const float* xy, int n, float rad, int periodic, int* counts, int* nbr, int cap) {
    float rad2 = rad * rad;
    for (int i = 0; i < n; i++) {
        int c = 0;
        for (int j = 0; j < n; j++) {
            if (i == j) continue;
            float dx = xy[2 * j] - xy[2 * i], dy = xy[2 * j + 1] - xy[2 * i + 1];
            if (periodic) { dx = min_image(dx); dy = min_image(dy); }
            float r2 = dx * dx + dy * dy;
            if (r2 <= rad2) {
                if (c < cap) nbr[(size_t)i * cap + c] = j;
                c++;
            }
        }
        counts[i] = c;
    }
    return 0;
}
