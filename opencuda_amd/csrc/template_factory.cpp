// template_factory.cpp -- the GAUSS and EUCLIDE template vendors and the
// pooled TemplateFactory (okano_3_0/TemplateFactory.cu).
//
// The weight formulas, the ring order of the spatial template, the pool
// hash, the 1e-6 sigma equality and the replacement policy are the
// reference's, because they decide which float bits the filter multiplies:
// a sigma within 1e-6 of a pooled one reuses that template's weights.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>

#include "TemplateFactory.h"
#include "opencuda_bf.h"

namespace {

// ---------------------------------------------------------------- vendors

// okano_3_0/TemplateFactory.cu:1057-1076
bool gauss_legal(TFSize size, const void *priv)
{
    if (priv == NULL || size.z != 1) return false;
    if (size.x < 3 || size.y != size.x || size.x % 2 == 0) return false;
    return std::fabs(*static_cast<const float *>(priv)) >= 1.0e-8f;
}

// okano_3_0/TemplateFactory.cu:1079-1089
int gauss_hash(TFSize size, const void *priv)
{
    if (priv == NULL) return -1;
    return ((int)(size.x / 2 - 1) ^ (int)(std::fabs(*static_cast<const float *>(priv)) * 10.0f)) %
           TF_VOL_SHAPE;
}

// Fill `offsets` (2 ints per point) and `weights` for an edge of `size`:
// (0,0) with weight 1, then rings i = 1..size/2, each as j = -i+1..i of
// (i,j), (-i,j), (j,i), (j,-i) sharing exp((float)(-(i*i+j*j)) / s22)
// (okano_3_0/TemplateFactory.cu:1092-1168).
int gauss_fill(float sigma, int size, int32_t *offsets, float *weights)
{
    const float sigma22 = 2 * sigma * sigma;
    int radius = size / 2;
    if (radius < 1) radius = 1;
    int n = 0;
    offsets[0] = 0;
    offsets[1] = 0;
    weights[n++] = 1.0f;
    for (int i = 1; i <= radius; i++) {
        for (int j = -i + 1; j <= i; j++) {
            const float w = std::exp((float)(-(i * i + j * j)) / sigma22);
            const int pts[4][2] = { { i, j }, { -i, j }, { j, i }, { j, -i } };
            for (const auto &p : pts) {
                offsets[2 * n] = p[0];
                offsets[2 * n + 1] = p[1];
                weights[n++] = w;
            }
        }
    }
    return n;
}

// okano_3_0/TemplateFactory.cu:1293-1310
bool euclid_legal(TFSize size, const void *priv)
{
    if (priv == NULL || size.z != 1) return false;
    if (size.x < 1 || (size.y != size.x && size.y != 1)) return false;
    return std::fabs(*static_cast<const float *>(priv)) >= 1.0e-8f;
}

// okano_3_0/TemplateFactory.cu:1313-1320
int euclid_hash(TFSize size, const void *priv)
{
    if (priv == NULL) return -1;
    return ((int)size.x ^ (int)(std::fabs(*static_cast<const float *>(priv)) * 10.0f)) %
           TF_VOL_SHAPE;
}

// okano_3_0/TemplateFactory.cu:1323-1373: point i = (i, 0), weight
// exp((float)(-(i*i)) / s22).
void euclid_fill(float sigma, int size, int32_t *offsets, float *weights)
{
    const float sigma22 = 2 * sigma * sigma;
    for (int i = 0; i < size; i++) {
        if (offsets) {
            offsets[2 * i] = i;
            offsets[2 * i + 1] = 0;
        }
        weights[i] = std::exp((float)(-(i * i)) / sigma22);
    }
}

Template *make_host_template(size_t count)
{
    Template *t = NULL;
    if (TemplateBasicOp::newTemplate(&t) != NO_ERROR) return NULL;
    if (TemplateBasicOp::makeAtHost(t, count) != NO_ERROR) {
        TemplateBasicOp::deleteTemplate(t);
        return NULL;
    }
    return t;
}

Template *gauss_create(TFSize size, const void *priv)
{
    const float sigma = *static_cast<const float *>(priv);
    if (std::fabs(sigma) < 1.0e-8f) return NULL;
    int radius = (int)size.x / 2;
    if (radius < 1) radius = 1;
    const int edge = 2 * radius + 1;
    Template *t = make_host_template((size_t)edge * edge);
    if (!t) return NULL;
    gauss_fill(sigma, edge, t->tplData, ATTACHED_DATA(t));
    return t;
}

Template *euclid_create(TFSize size, const void *priv)
{
    const float sigma = *static_cast<const float *>(priv);
    if (size.x < 1 || std::fabs(sigma) < 1.0e-8f) return NULL;
    Template *t = make_host_template(size.x);
    if (!t) return NULL;
    euclid_fill(sigma, (int)size.x, t->tplData, ATTACHED_DATA(t));
    return t;
}

// okano_3_0/TemplateFactory.cu:1171-1188 (gauss) and :1376-1393 (euclide)
bool sigma_equal(TFSize a, const void *pa, TFSize b, const void *pb)
{
    if (pa == NULL || pb == NULL) return false;
    if (std::fabs(*static_cast<const float *>(pa) - *static_cast<const float *>(pb)) >= 1.0e-6f)
        return false;
    return a.x == b.x;
}

// ------------------------------------------------------------------- pool

constexpr int kSlots = TF_VOL_SHAPE * TF_SET_SHAPE;

struct Slot {
    Template *tpl = NULL;
    TFSize size{ 0, 0, 0 };
    float sigma = 0.0f;   // private parameter copy
    bool has_sigma = false;
    int count = 0;        // outstanding getTemplate references
};

struct Pool {
    std::mutex mu;
    Slot slot[TF_CNT_SHAPE][kSlots];
};

Pool &pool()
{
    static Pool *p = new Pool();   // never destroyed: templates may outlive statics
    return *p;
}

// Swap an entry with its predecessor in the set (okano_3_0/TemplateFactory.cu
// :1449-1502) so that frequently used templates are kicked last.
void boost(int shape, int idx)
{
    if (idx % TF_SET_SHAPE == 0) return;
    Slot tmp = pool().slot[shape][idx];
    pool().slot[shape][idx] = pool().slot[shape][idx - 1];
    pool().slot[shape][idx - 1] = tmp;
}

}  // namespace

// okano_3_0/TemplateFactory.cu:1505-1650
int TemplateFactory::getTemplate(Template **tpl, int shape, TFSize size, void *privated)
{
    if (shape < 0 || shape >= TF_CNT_SHAPE) return INVALID_DATA;
    if (tpl == NULL) return NULL_POINTER;
    if (shape != TF_SHAPE_GAUSS && shape != TF_SHAPE_EUCLIDE) return UNIMPLEMENT;
    const bool gauss = shape == TF_SHAPE_GAUSS;
    if (!(gauss ? gauss_legal(size, privated) : euclid_legal(size, privated))) return INVALID_DATA;
    const int hash = gauss ? gauss_hash(size, privated) : euclid_hash(size, privated);
    if (hash < 0 || hash > TF_VOL_SHAPE) return INVALID_DATA;
    std::lock_guard<std::mutex> lock(pool().mu);
    Slot *slots = pool().slot[shape];
    const int start = hash * TF_SET_SHAPE, end = start + TF_SET_SHAPE;
    int pos;
    for (pos = start; pos < end; pos++) {
        if (slots[pos].tpl == NULL) break;
        if (sigma_equal(slots[pos].size, slots[pos].has_sigma ? &slots[pos].sigma : NULL, size,
                        privated)) {
            *tpl = slots[pos].tpl;
            slots[pos].count++;
            boost(shape, pos);
            return NO_ERROR;
        }
    }
    *tpl = gauss ? gauss_create(size, privated) : euclid_create(size, privated);
    if (*tpl == NULL) return OUT_OF_MEM;
    if (pos >= end) {
        // Set full: replace the last unused entry in the back half.
        bool replaced = false;
        pos = end - 1;
        const int stop = (start + end) / 2;
        do {
            if (slots[pos].count == 0) {
                replaced = true;
                TemplateBasicOp::deleteTemplate(slots[pos].tpl);
                break;
            }
        } while (pos-- >= stop);
        if (!replaced) pos = -1;
    }
    if (pos >= start && pos < end) {
        slots[pos].tpl = *tpl;
        slots[pos].size = size;
        slots[pos].sigma = *static_cast<const float *>(privated);
        slots[pos].has_sigma = true;
        slots[pos].count = 1;
    }
    return NO_ERROR;
}

// okano_3_0/TemplateFactory.cu:1653-1675
int TemplateFactory::putTemplate(Template *tpl)
{
    if (tpl == NULL) return NULL_POINTER;
    {
        std::lock_guard<std::mutex> lock(pool().mu);
        for (int shape = 0; shape < TF_CNT_SHAPE; shape++) {
            for (int i = 0; i < kSlots; i++) {
                Slot &s = pool().slot[shape][i];
                if (s.tpl == tpl) {
                    if (s.count > 0) s.count--;
                    return NO_ERROR;
                }
            }
        }
    }
    TemplateBasicOp::deleteTemplate(tpl);
    return NO_ERROR;
}

extern "C" int bf_gauss_template(float sigma, int32_t size, int32_t *offsets, float *weights)
{
    if (!offsets || !weights) return NULL_POINTER;
    TFSize sz((unsigned)size, (unsigned)size);
    if (size < 3 || !gauss_legal(sz, &sigma)) return INVALID_DATA;
    return gauss_fill(sigma, size, offsets, weights);
}

extern "C" int bf_euclid_template(float sigma, int32_t size, float *weights)
{
    if (!weights) return NULL_POINTER;
    TFSize sz((unsigned)size, 1);
    if (size < 1 || !euclid_legal(sz, &sigma)) return INVALID_DATA;
    euclid_fill(sigma, size, NULL, weights);
    return size;
}
