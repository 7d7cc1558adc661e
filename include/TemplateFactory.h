// TemplateFactory.h -- pooled template factory.
//
// Same interface as okano_3_0/TemplateFactory.h:70-148, all five shapes:
//   TF_SHAPE_BOX      size.x by size.y box as (x, y) pairs, weights 1/(x*y),
//                     no private parameter (okano_3_0/TemplateFactory.cu
//                     :340-402; like the reference, only (x-1)(y-1) of the
//                     x*y points are box points -- its loops are half-open --
//                     and the remaining entries, unwritten there, are the
//                     centre with weight 0 here);
//   TF_SHAPE_CIRCLE   filled digital disc of diameter size.x (odd, >= 3,
//                     size.y == size.x), weights 1/count (:501-666);
//   TF_SHAPE_ARC      ring between the circles of diameters size.x < size.y
//                     (both odd), weights 1/count (:771-936);
//   TF_SHAPE_GAUSS    ring-ordered spatial template with Gaussian weights,
//                     size = edge (odd, >= 3), private parameter = float sigma
//                     (okano_3_0/TemplateFactory.cu:1057-1168);
//   TF_SHAPE_EUCLIDE  range table exp(-d^2 / (2 sigma^2)), d < size,
//                     private parameter = float sigma (:1293-1373).
// Templates are pooled per shape in
// TF_VOL_SHAPE sets of TF_SET_SHAPE entries; two requests whose sigmas
// differ by less than 1e-6 share one template (:1171-1188).  getTemplate and
// putTemplate are serialised by a mutex.
#ifndef __TEMPLATEFACTORY_H__
#define __TEMPLATEFACTORY_H__

#include <stddef.h>

#include "ErrorCode.h"
#include "Template.h"

#define TF_CNT_SHAPE       5
#define TF_VOL_SHAPE     128
#define TF_SET_SHAPE       4

#define TF_SHAPE_BOX       0
#define TF_SHAPE_CIRCLE    1
#define TF_SHAPE_ARC       2
#define TF_SHAPE_GAUSS     3
#define TF_SHAPE_EUCLIDE   4

// Template size (the reference passes a CUDA dim3; x is the edge or length).
struct TFSize {
    unsigned int x, y, z;
    TFSize(unsigned int vx = 1, unsigned int vy = 1, unsigned int vz = 1)
        : x(vx), y(vy), z(vz) {}
};

class TemplateFactory {
public:
    // Square size `size` x `size` (okano_3_0/TemplateFactory.h:113-123).
    static int getTemplate(Template **tpl, int shape, size_t size, void *privated = NULL)
    {
        return getTemplate(tpl, shape, TFSize((unsigned)size, (unsigned)size), privated);
    }
    static int getTemplate(Template **tpl, int shape, TFSize size, void *privated = NULL);
    // Return a template from getTemplate: pooled ones are released, others freed.
    static int putTemplate(Template *tpl);
};

#endif
