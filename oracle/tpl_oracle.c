/*
 * tpl_oracle.c -- CPU restatement of the BOX, CIRCLE and ARC template
 * vendors of okano_3_0/TemplateFactory.cu (the shapes the sibling filters
 * of the bilateral filter use).
 *
 * TEST INFRASTRUCTURE ONLY, like bf_oracle.c: only tests/ may load the
 * library built from it; the product path never links or calls it.
 *
 * orc_box_template     :363-402  x*y entries; points (c, r) for c in
 *                      [-(x-1)/2, x/2), r in [-(y-1)/2, y/2), row-major
 *                      (the reference's loops are half-open, so only
 *                      (x-1)(y-1) entries are box points; the rest are not
 *                      written there -- returned here as the centre with
 *                      weight 0); weights 1/(x*y).
 * orc_circle_template  :527-666  octant midpoint walk from (0, r), filling
 *                      rows between symmetric points; weights 1/count.
 * orc_arc_template     :798-936  two octant walks (inner radius from
 *                      size.x, outer from size.y), points between them
 *                      with their 8 (4 on the diagonal) symmetric images.
 * Points are (x, y) pairs; each function returns the count, or -1 for an
 * illegal size; with pts == NULL it only counts.
 */
#include <stdint.h>
#include <stdlib.h>

static void put(int32_t *pts, int *n, int a, int b)
{
    if (pts) {
        pts[2 * *n] = a;
        pts[2 * *n + 1] = b;
    }
    (*n)++;
}

int orc_box_template(int sx, int sy, int32_t *pts, float *w)
{
    if (sx < 1 || sy < 1) return -1;
    const int count = sx * sy;
    if (!pts || !w) return count;
    int n = 0;
    for (int r = -((sy - 1) / 2); r < sy / 2; r++)
        for (int c = -((sx - 1) / 2); c < sx / 2; c++) {
            w[n] = 1.0f / count;
            put(pts, &n, c, r);
        }
    for (; n < count; n++) {
        pts[2 * n] = 0;
        pts[2 * n + 1] = 0;
        w[n] = 0.0f;
    }
    return count;
}

static void fill_weights(float *w, int n)
{
    if (!w) return;
    const float a = 1.0f / n;
    for (int i = 0; i < n; i++) w[i] = a;
}

int orc_circle_template(int size, int32_t *pts, float *w)
{
    if (size < 3 || size % 2 == 0) return -1;
    int radius = (size - 1) / 2;
    if (radius < 1) radius = 1;
    const int r2 = radius * radius;
    int n = 0;
    put(pts, &n, 0, radius);
    put(pts, &n, 0, -radius);
    for (int ix = -radius; ix <= radius; ix++) put(pts, &n, ix, 0);
    int x = 0, y = radius;
    while (x < y) {
        x++;
        int d1 = abs(x * x + y * y - r2);
        int d2 = abs(x * x + (y - 1) * (y - 1) - r2);
        if (d1 < d2) {
            put(pts, &n, x, y);
            put(pts, &n, -x, y);
            put(pts, &n, x, -y);
            put(pts, &n, -x, -y);
        } else {
            y--;
            if (x > y) break;
            for (int ix = -x; ix <= x; ix++) {
                put(pts, &n, ix, y);
                put(pts, &n, ix, -y);
            }
        }
        if (x < y)
            for (int iy = -y; iy <= y; iy++) {
                put(pts, &n, iy, x);
                put(pts, &n, iy, -x);
            }
    }
    if (pts) fill_weights(w, n);
    return n;
}

int orc_arc_template(int sx, int sy, int32_t *pts, float *w)
{
    if (sx < 1 || sy <= sx || sx % 2 == 0 || sy % 2 == 0) return -1;
    int sr = (sx - 1) / 2;
    sr = sr < 1 ? 1 : sr;
    const int sr2 = sr * sr;
    int lr = (sy - 1) / 2;
    lr = lr <= sr ? sr + 1 : lr;
    const int lr2 = lr * lr;
    int n = 0;
    for (int y = sr; y < lr; y++) {
        put(pts, &n, 0, y);
        put(pts, &n, 0, -y);
        put(pts, &n, y, 0);
        put(pts, &n, -y, 0);
    }
    int x = 0, iy = sr, oy = lr;
    while (x < oy) {
        x++;
        if (x - 1 < iy) {
            const int d1 = abs(x * x + iy * iy - sr2);
            const int d2 = abs(x * x + (iy - 1) * (iy - 1) - sr2);
            if (d1 > d2) iy--;
        }
        if (x >= iy) iy = x;
        const int e1 = abs(x * x + oy * oy - lr2);
        const int e2 = abs(x * x + (oy - 1) * (oy - 1) - lr2);
        if (e1 > e2) oy--;
        if (x > oy) break;
        for (int y = iy; y < oy; y++) {
            put(pts, &n, x, y);
            put(pts, &n, -x, y);
            put(pts, &n, x, -y);
            put(pts, &n, -x, -y);
            if (x == y) continue;
            put(pts, &n, y, x);
            put(pts, &n, -y, x);
            put(pts, &n, y, -x);
            put(pts, &n, -y, -x);
        }
    }
    if (pts) fill_weights(w, n);
    return n;
}
