// Template.h -- sparse point template with attached weights.
//
// The reference's type (okano_3_0/Template.h:14-46): `count` points stored as
// (first, second) int pairs in tplData, plus one float of attached data per
// point.  The bilateral filter reads the first coordinate as the row offset
// and the second as the column offset (okano_3_0/BilateralFilter.cu:116-120).
// deviceId < 0 means host memory, else the device holding both arrays.
#ifndef __TEMPLATE_H__
#define __TEMPLATE_H__

#include <stddef.h>

#include "ErrorCode.h"

typedef struct Template_st {
    size_t count;   // number of points
    int *tplData;   // 2*count coordinates, point i = (tplData[2i], tplData[2i+1])
} Template;

typedef struct TemplateCuda_st {
    Template tplMeta;     // the public part
    float *attachedData;  // count weights
    int deviceId;         // < 0: host memory
} TemplateCuda;

#define TEMPLATE_CUDA(tpl) \
    ((TemplateCuda *)((unsigned char *)(tpl) - offsetof(TemplateCuda, tplMeta)))
#define ATTACHED_DATA(tpl) (TEMPLATE_CUDA(tpl)->attachedData)

class TemplateBasicOp {
public:
    static int newTemplate(Template **outtpl);
    static int deleteTemplate(Template *intpl);
    static int makeAtCurrentDevice(Template *tpl, size_t count);
    static int makeAtHost(Template *tpl, size_t count);
    // TPLT file: "TPLT", u32 count, 20 reserved bytes, 2*count int32
    // coordinates, count float32 weights (okano_3_0/Template.cu:174-382).
    static int readFromFile(const char *filepath, Template *outtpl);
    static int writeToFile(const char *filepath, Template *intpl);
    static int copyToCurrentDevice(Template *tpl);
    static int copyToCurrentDevice(Template *srctpl, Template *dsttpl);
    static int copyToHost(Template *tpl);
    static int copyToHost(Template *srctpl, Template *dsttpl);
};

#endif
