/*
 * region_oracle.c -- CPU restatement of the two stages around the bilateral
 * filter in okano_3_0/SimpleRegionDetect.cu:287-312 (LocalCluster before it,
 * ConnectRegion after it).
 *
 * TEST INFRASTRUCTURE ONLY, like bf_oracle.c: only tests/ may load the
 * library built from it; the product path never links or calls it.
 *
 * orc_local_cluster    okano_3_0/LocalCluster.cu:45-153 (kernel), :246-270
 *                      (guard band sx/ex/sy/ey, trimming), defaults
 *                      LocalCluster.h.  Per pixel and per direction z of the
 *                      eight in _idxDev (:14-19): walk while the gray gap and
 *                      the side-sum difference stay within their thresholds,
 *                      average what was walked; then replace the 8 - pntcount
 *                      averages farthest from the centre by the centre and
 *                      average.  Every read of the reference is a linear
 *                      address data + y * pitch + x from the ROI origin; the
 *                      walks reach 2 (pntRange - 1) pixels from the centre
 *                      while the guard band is only pntRange + 2 wide, so
 *                      near the image border the reference reads outside the
 *                      ROI and possibly outside the allocation.  Here reads
 *                      use the same linear address, clamped into the
 *                      allocation [lo, hi) (byte offsets from the ROI
 *                      origin): identical to the reference wherever its reads
 *                      stay inside the allocation, defined where they do not.
 *
 * orc_connect_region   okano_3_0/ConnectRegion.cu:789-1043 (host flow) and
 *                      its kernels :184-787.  Two 8-neighbours are connected
 *                      iff both gray values lie in [mingray, maxgray] (the
 *                      CONNREGION_DIFF_INT macro, :53; the class's threshold
 *                      member is never used); labels converge to the minimum
 *                      raster index of each component (_findRootDev and the
 *                      atomicMin unions); areas are counted per root by
 *                      _computeAreaKer (:491-593), whose launch covers only
 *                      rows below 256 * ceil(floor((H + 4) / 32) / 8)
 *                      (CONNREGION_PACK_MASK is 4, not 31, :45); a root with
 *                      a counted area outside [minArea, maxArea] is dropped
 *                      (:595-617, area 0 is never dropped); kept roots are
 *                      numbered in raster order (:656-758), numbers >= 256
 *                      dropped; output = number + 1 as u8, 0 for dropped.
 *
 * Built with -ffp-contract=off and without fast-math like bf_oracle.c.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* _idxDev (LocalCluster.cu:14-19): {dx, dy} for z = 0..7 */
static const int lc_dir[8][2] = {
    { 1, 0 }, { 1, -1 }, { 0, -1 }, { -1, -1 }, { -1, 0 }, { -1, 1 }, { 0, 1 }, { 1, 1 }
};

static inline uint8_t lc_read(const uint8_t *base, int64_t pitch, int64_t lo, int64_t hi, int y, int x)
{
    int64_t off = (int64_t)y * pitch + x;
    if (off < lo) off = lo;
    if (off >= hi) off = hi - 1;
    return base[off];
}

static inline uint8_t sat_u8(float q)
{
    if (!(q > 0.0f)) return 0;
    if (q >= 255.0f) return 255;
    return (uint8_t)(int)q;
}

int orc_local_cluster(const uint8_t *base, int64_t pitch, int64_t lo, int64_t hi, int W, int H,
                      uint8_t *out, int64_t outpitch, int pntrange, int gapthred, int diffethred,
                      int problack, int prowhite, int pntcount)
{
    if (!base || !out || W <= 0 || H <= 0 || lo > 0 || hi <= 0) return -1;
    const int sx = pntrange + 2, ex = W - pntrange - 2;       /* LocalCluster.cu:254-257 */
    const int sy = pntrange + 2, ey = H - pntrange - 2;
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            const uint8_t cv = lc_read(base, pitch, lo, hi, y, x);
            if (x < sx || x > ex || y < sy || y > ey || cv <= problack || cv >= prowhite) {
                out[(int64_t)y * outpitch + x] = cv;
                continue;
            }
            float temp[8];
            for (int z = 0; z < 8; z++) {
                const int ddx = lc_dir[z][0], ddy = lc_dir[z][1];
                uint8_t cuv = lc_read(base, pitch, lo, hi, y + ddy, x + ddx);
                uint8_t pre1 = lc_read(base, pitch, lo, hi, y - ddy, x - ddx);
                uint8_t pre2 = cv;
                uint8_t latt1 = lc_read(base, pitch, lo, hi, y + 2 * ddy, x + 2 * ddx);
                float sum = (float)cv;
                int i, px = x, py = y;
                for (i = 1; i < pntrange; i++) {
                    if (fabsf((float)cv - (float)cuv) > (float)gapthred) break;
                    px += 2 * ddx;
                    py += 2 * ddy;
                    const uint8_t latt2 = lc_read(base, pitch, lo, hi, py, px);
                    const float side1 = (float)(pre1 + pre2);
                    const float side2 = (float)(latt1 + latt2);
                    if (fabsf(side1 - side2) > (float)diffethred) break;
                    pre1 = pre2;
                    pre2 = cuv;
                    cuv = latt1;
                    latt1 = latt2;
                    sum += (float)cuv;
                }
                temp[z] = sum / (float)i;
            }
            /* z == 0 thread's selection (LocalCluster.cu:132-152) */
            int mark = 0;
            for (int j = 8 - pntcount; j > 0; j--) {
                for (int i = 0; i < 8; i++)
                    if (fabsf(temp[i] - (float)cv) >= fabsf(temp[mark] - (float)cv)) mark = i;
                temp[mark] = (float)cv;
            }
            float sumag = 0.0f;
            for (int j = 0; j < 8; j++) sumag += temp[j];
            sumag -= (float)((8 - pntcount) * cv);
            const float q = sumag / (float)pntcount;
            out[(int64_t)y * outpitch + x] = q > 255.0f ? 255 : sat_u8(q);
        }
    }
    return 0;
}

static int cr_find(int *p, int x)
{
    int r = x;
    while (p[r] != r) r = p[r];
    while (p[x] != r) {           /* path compression */
        const int n = p[x];
        p[x] = r;
        x = n;
    }
    return r;
}

static void cr_union(int *p, int a, int b)
{
    a = cr_find(p, a);
    b = cr_find(p, b);
    if (a < b) p[b] = a;
    else if (b < a) p[a] = b;
}

int orc_connect_region(const uint8_t *in, int64_t inpitch, int W, int H, uint8_t *out,
                       int64_t outpitch, int min_area, int max_area, int mingray, int maxgray)
{
    if (!in || !out || W <= 0 || H <= 0) return -1;
    const int64_t n = (int64_t)W * H;
    int *p = (int *)malloc((size_t)n * sizeof(int));
    int *area = (int *)calloc((size_t)n, sizeof(int));
    int *seq = (int *)malloc((size_t)n * sizeof(int));
    if (!p || !area || !seq) {
        free(p);
        free(area);
        free(seq);
        return -6;
    }
#define FG(v) ((v) >= mingray && (v) <= maxgray)
    for (int64_t i = 0; i < n; i++) p[i] = (int)i;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const uint8_t v = in[(int64_t)y * inpitch + x];
            if (!FG(v)) continue;
            const int idx = y * W + x;
            /* the four neighbours before (x, y) in raster order cover all 8 */
            static const int nb[4][2] = { { -1, 0 }, { -1, -1 }, { 0, -1 }, { 1, -1 } };
            for (int k = 0; k < 4; k++) {
                const int xx = x + nb[k][0], yy = y + nb[k][1];
                if (xx < 0 || xx >= W || yy < 0) continue;
                if (FG(in[(int64_t)yy * inpitch + xx])) cr_union(p, idx, yy * W + xx);
            }
        }
#undef FG
    for (int64_t i = 0; i < n; i++) p[i] = cr_find(p, (int)i);
    /* rows the reference's area launch covers */
    const int packed = (H + 4) / 32;
    int covered = 256 * ((packed + 7) / 8);
    if (covered > H) covered = H;
    for (int64_t i = 0; i < (int64_t)covered * W; i++) area[p[i]]++;
    int next = 0;
    for (int64_t i = 0; i < n; i++) {
        seq[i] = -1;
        if (p[i] != i) continue;
        const int a = area[i];
        if (a != 0 && (a < min_area || a > max_area)) continue;
        seq[i] = next < 256 ? next : -1;
        next++;
    }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const int s = seq[p[y * W + x]];
            out[(int64_t)y * outpitch + x] = (uint8_t)(s + 1);
        }
    free(p);
    free(area);
    free(seq);
    return 0;
}
