// template.cpp -- TemplateBasicOp over the HIP runtime.
//
// Behaviour and status codes follow okano_3_0/Template.cu (cited per
// method).  Deleting a template also drops any device filter tables the
// bilateral filter built from it.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <new>

#include "Template.h"

namespace bfimpl {
void lut_cache_forget(const Template *tpl);
}

namespace {

int device_count()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int current_device()
{
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) return 0;
    return d;
}

void free_arrays(int *xy, float *att, int dev)
{
    if (dev < 0) {
        delete[] xy;
        delete[] att;
        return;
    }
    const int cur = current_device();
    if (dev != cur) hipSetDevice(dev);
    hipFree(xy);
    hipFree(att);
    if (dev != cur) hipSetDevice(cur);
}

}  // namespace

// okano_3_0/Template.cu:14-35
int TemplateBasicOp::newTemplate(Template **outtpl)
{
    if (outtpl == NULL) return NULL_POINTER;
    TemplateCuda *c = new (std::nothrow) TemplateCuda;
    if (!c) return OUT_OF_MEM;
    c->tplMeta.count = 0;
    c->tplMeta.tplData = NULL;
    c->attachedData = NULL;
    c->deviceId = -1;
    *outtpl = &c->tplMeta;
    return NO_ERROR;
}

// okano_3_0/Template.cu:38-82
int TemplateBasicOp::deleteTemplate(Template *intpl)
{
    if (intpl == NULL) return NULL_POINTER;
    TemplateCuda *c = TEMPLATE_CUDA(intpl);
    if (c->deviceId >= device_count()) return OP_OVERFLOW;
    bfimpl::lut_cache_forget(intpl);
    free_arrays(intpl->tplData, c->attachedData, c->deviceId);
    delete c;
    return NO_ERROR;
}

// okano_3_0/Template.cu:85-132
int TemplateBasicOp::makeAtCurrentDevice(Template *tpl, size_t count)
{
    if (tpl == NULL) return NULL_POINTER;
    if (count < 1) return INVALID_DATA;
    if (tpl->tplData != NULL) return UNMATCH_IMG;
    TemplateCuda *c = TEMPLATE_CUDA(tpl);
    void *xy = NULL, *att = NULL;
    if (hipMalloc(&xy, 2 * count * sizeof(int)) != hipSuccess) {
        tpl->tplData = NULL;
        return CUDA_ERROR;
    }
    if (hipMalloc(&att, count * sizeof(float)) != hipSuccess) {
        hipFree(xy);
        tpl->tplData = NULL;
        c->attachedData = NULL;
        return CUDA_ERROR;
    }
    tpl->tplData = static_cast<int *>(xy);
    c->attachedData = static_cast<float *>(att);
    tpl->count = count;
    c->deviceId = current_device();
    return NO_ERROR;
}

// okano_3_0/Template.cu:135-171
int TemplateBasicOp::makeAtHost(Template *tpl, size_t count)
{
    if (tpl == NULL) return NULL_POINTER;
    if (count < 1) return INVALID_DATA;
    if (tpl->tplData != NULL) return UNMATCH_IMG;
    TemplateCuda *c = TEMPLATE_CUDA(tpl);
    tpl->tplData = new (std::nothrow) int[count * 2];
    if (tpl->tplData == NULL) return OUT_OF_MEM;
    c->attachedData = new (std::nothrow) float[count];
    if (c->attachedData == NULL) {
        delete[] tpl->tplData;
        tpl->tplData = NULL;
        return OUT_OF_MEM;
    }
    tpl->count = count;
    c->deviceId = -1;
    return NO_ERROR;
}

// okano_3_0/Template.cu:174-295
int TemplateBasicOp::readFromFile(const char *filepath, Template *outtpl)
{
    if (filepath == NULL || outtpl == NULL) return NULL_POINTER;
    TemplateCuda *c = TEMPLATE_CUDA(outtpl);
    if (c->deviceId >= device_count()) return OP_OVERFLOW;
    std::ifstream f(filepath, std::ios::in | std::ios::binary);
    if (!f) return NO_FILE;
    char magic[4] = { 0, 0, 0, 0 };
    f.read(magic, 4);
    if (!f || std::memcmp(magic, "TPLT", 4) != 0) return WRONG_FILE;
    uint32_t count32 = 0;
    f.read(reinterpret_cast<char *>(&count32), 4);
    if (!f || count32 < 1) return WRONG_FILE;
    const size_t count = count32;
    char reserved[20];
    f.read(reserved, 20);
    int *xy;
    float *att;
    bool reuse;
    if (outtpl->tplData != NULL && outtpl->count == count && c->deviceId == -1) {
        xy = outtpl->tplData;
        att = c->attachedData;
        reuse = true;
    } else {
        xy = new (std::nothrow) int[count * 2];
        att = new (std::nothrow) float[count];
        reuse = false;
        if (xy == NULL || att == NULL) {
            delete[] xy;
            delete[] att;
            return OUT_OF_MEM;
        }
    }
    static_assert(sizeof(int) == 4, "TPLT stores 32-bit coordinates");
    f.read(reinterpret_cast<char *>(xy), (std::streamsize)(count * 2 * 4));
    f.read(reinterpret_cast<char *>(att), (std::streamsize)(count * 4));
    if (!f) {
        if (!reuse) {
            delete[] xy;
            delete[] att;
        }
        return WRONG_FILE;
    }
    if (outtpl->tplData != NULL && !reuse) {
        bfimpl::lut_cache_forget(outtpl);
        free_arrays(outtpl->tplData, c->attachedData, c->deviceId);
    } else if (reuse) {
        bfimpl::lut_cache_forget(outtpl);
    }
    outtpl->count = count;
    outtpl->tplData = xy;
    c->attachedData = att;
    c->deviceId = -1;
    return NO_ERROR;
}

// okano_3_0/Template.cu:298-382
int TemplateBasicOp::writeToFile(const char *filepath, Template *intpl)
{
    if (filepath == NULL || intpl == NULL) return NULL_POINTER;
    std::ofstream f(filepath, std::ios::out | std::ios::binary);
    if (!f) return NO_FILE;
    int err = copyToHost(intpl);
    if (err < 0) return err;
    TemplateCuda *c = TEMPLATE_CUDA(intpl);
    const uint32_t count32 = (uint32_t)intpl->count;
    const char reserved[20] = { 0 };
    f.write("TPLT", 4);
    f.write(reinterpret_cast<const char *>(&count32), 4);
    f.write(reserved, 20);
    f.write(reinterpret_cast<const char *>(intpl->tplData),
            (std::streamsize)(intpl->count * 2 * 4));
    f.write(reinterpret_cast<const char *>(c->attachedData), (std::streamsize)(intpl->count * 4));
    return f ? NO_ERROR : NO_FILE;
}

// okano_3_0/Template.cu:385-500
int TemplateBasicOp::copyToCurrentDevice(Template *tpl)
{
    if (tpl == NULL) return NULL_POINTER;
    TemplateCuda *c = TEMPLATE_CUDA(tpl);
    if (c->deviceId >= device_count()) return OP_OVERFLOW;
    const int cur = current_device();
    if (tpl->tplData == NULL || tpl->count == 0) return UNMATCH_IMG;
    if (c->deviceId == cur) return NO_ERROR;
    void *xy = NULL, *att = NULL;
    if (hipMalloc(&xy, tpl->count * 2 * sizeof(int)) != hipSuccess) return CUDA_ERROR;
    if (hipMalloc(&att, tpl->count * sizeof(float)) != hipSuccess) {
        hipFree(xy);
        return CUDA_ERROR;
    }
    hipError_t e1, e2;
    if (c->deviceId < 0) {
        e1 = hipMemcpy(xy, tpl->tplData, tpl->count * 2 * sizeof(int), hipMemcpyHostToDevice);
        e2 = hipMemcpy(att, c->attachedData, tpl->count * sizeof(float), hipMemcpyHostToDevice);
    } else {
        e1 = hipMemcpyPeer(xy, cur, tpl->tplData, c->deviceId, tpl->count * 2 * sizeof(int));
        e2 = hipMemcpyPeer(att, cur, c->attachedData, c->deviceId, tpl->count * sizeof(float));
    }
    if (e1 != hipSuccess || e2 != hipSuccess) {
        hipFree(xy);
        hipFree(att);
        return CUDA_ERROR;
    }
    free_arrays(tpl->tplData, c->attachedData, c->deviceId);
    tpl->tplData = static_cast<int *>(xy);
    c->attachedData = static_cast<float *>(att);
    c->deviceId = cur;
    return NO_ERROR;
}

// okano_3_0/Template.cu:503-640
int TemplateBasicOp::copyToCurrentDevice(Template *srctpl, Template *dsttpl)
{
    if (srctpl == NULL) return NULL_POINTER;
    if (dsttpl == NULL || dsttpl == srctpl) return copyToCurrentDevice(srctpl);
    TemplateCuda *s = TEMPLATE_CUDA(srctpl);
    TemplateCuda *d = TEMPLATE_CUDA(dsttpl);
    if (srctpl->tplData == NULL || srctpl->count == 0) return INVALID_DATA;
    const int ndev = device_count();
    if (s->deviceId >= ndev || d->deviceId >= ndev) return OP_OVERFLOW;
    const int cur = current_device();
    const size_t n = srctpl->count;
    bool reuse = dsttpl->tplData != NULL && d->deviceId == cur && dsttpl->count == n;
    int *xy = reuse ? dsttpl->tplData : NULL;
    float *att = reuse ? d->attachedData : NULL;
    if (!reuse) {
        void *a = NULL, *b = NULL;
        if (hipMalloc(&a, n * 2 * sizeof(int)) != hipSuccess) return CUDA_ERROR;
        if (hipMalloc(&b, n * sizeof(float)) != hipSuccess) {
            hipFree(a);
            return CUDA_ERROR;
        }
        xy = static_cast<int *>(a);
        att = static_cast<float *>(b);
    }
    hipError_t e1, e2;
    if (s->deviceId < 0) {
        e1 = hipMemcpy(xy, srctpl->tplData, n * 2 * sizeof(int), hipMemcpyHostToDevice);
        e2 = hipMemcpy(att, s->attachedData, n * sizeof(float), hipMemcpyHostToDevice);
    } else {
        e1 = hipMemcpyPeer(xy, cur, srctpl->tplData, s->deviceId, n * 2 * sizeof(int));
        e2 = hipMemcpyPeer(att, cur, s->attachedData, s->deviceId, n * sizeof(float));
    }
    if (e1 != hipSuccess || e2 != hipSuccess) {
        if (!reuse) {
            hipFree(xy);
            hipFree(att);
        }
        return CUDA_ERROR;
    }
    bfimpl::lut_cache_forget(dsttpl);
    if (!reuse && dsttpl->tplData != NULL) free_arrays(dsttpl->tplData, d->attachedData, d->deviceId);
    dsttpl->count = n;
    dsttpl->tplData = xy;
    d->attachedData = att;
    d->deviceId = cur;
    return NO_ERROR;
}

// okano_3_0/Template.cu:643-720
int TemplateBasicOp::copyToHost(Template *tpl)
{
    if (tpl == NULL) return NULL_POINTER;
    TemplateCuda *c = TEMPLATE_CUDA(tpl);
    if (c->deviceId >= device_count()) return OP_OVERFLOW;
    if (tpl->tplData == NULL || tpl->count == 0) return UNMATCH_IMG;
    if (c->deviceId < 0) return NO_ERROR;
    int *xy = new (std::nothrow) int[tpl->count * 2];
    float *att = new (std::nothrow) float[tpl->count];
    if (!xy || !att) {
        delete[] xy;
        delete[] att;
        return OUT_OF_MEM;
    }
    const int cur = current_device();
    if (c->deviceId != cur) hipSetDevice(c->deviceId);
    hipError_t e1 = hipMemcpy(xy, tpl->tplData, tpl->count * 2 * sizeof(int), hipMemcpyDeviceToHost);
    hipError_t e2 = hipMemcpy(att, c->attachedData, tpl->count * sizeof(float), hipMemcpyDeviceToHost);
    if (c->deviceId != cur) hipSetDevice(cur);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        delete[] xy;
        delete[] att;
        return CUDA_ERROR;
    }
    free_arrays(tpl->tplData, c->attachedData, c->deviceId);
    tpl->tplData = xy;
    c->attachedData = att;
    c->deviceId = -1;
    return NO_ERROR;
}

// okano_3_0/Template.cu:723-860
int TemplateBasicOp::copyToHost(Template *srctpl, Template *dsttpl)
{
    if (srctpl == NULL) return NULL_POINTER;
    if (dsttpl == NULL || dsttpl == srctpl) return copyToHost(srctpl);
    TemplateCuda *s = TEMPLATE_CUDA(srctpl);
    TemplateCuda *d = TEMPLATE_CUDA(dsttpl);
    if (srctpl->tplData == NULL || srctpl->count == 0) return INVALID_DATA;
    const int ndev = device_count();
    if (s->deviceId >= ndev || d->deviceId >= ndev) return OP_OVERFLOW;
    const size_t n = srctpl->count;
    bool reuse = dsttpl->tplData != NULL && d->deviceId < 0 && dsttpl->count == n;
    int *xy = reuse ? dsttpl->tplData : new (std::nothrow) int[n * 2];
    float *att = reuse ? d->attachedData : new (std::nothrow) float[n];
    if (!xy || !att) {
        if (!reuse) {
            delete[] xy;
            delete[] att;
        }
        return OUT_OF_MEM;
    }
    if (s->deviceId < 0) {
        std::memcpy(xy, srctpl->tplData, n * 2 * sizeof(int));
        std::memcpy(att, s->attachedData, n * sizeof(float));
    } else {
        const int cur = current_device();
        if (s->deviceId != cur) hipSetDevice(s->deviceId);
        hipError_t e1 = hipMemcpy(xy, srctpl->tplData, n * 2 * sizeof(int), hipMemcpyDeviceToHost);
        hipError_t e2 = hipMemcpy(att, s->attachedData, n * sizeof(float), hipMemcpyDeviceToHost);
        if (s->deviceId != cur) hipSetDevice(cur);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            if (!reuse) {
                delete[] xy;
                delete[] att;
            }
            return CUDA_ERROR;
        }
    }
    bfimpl::lut_cache_forget(dsttpl);
    if (!reuse && dsttpl->tplData != NULL) free_arrays(dsttpl->tplData, d->attachedData, d->deviceId);
    dsttpl->count = n;
    dsttpl->tplData = xy;
    d->attachedData = att;
    d->deviceId = -1;
    return NO_ERROR;
}
