// image.cpp -- ImageBasicOp over the HIP runtime.
//
// Behaviour and status codes follow okano_3_0/Image.cu (cited per method):
// device images are pitched (hipMallocPitch), host images dense; in-place
// moves free the old copy; out-of-place copies reuse the destination's
// storage when its shape and residency already match.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <new>

#include "Image.h"

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

void free_data(unsigned char *p, int dev)
{
    if (!p) return;
    if (dev < 0) {
        delete[] p;
    } else {
        int cur = current_device();
        if (dev != cur) hipSetDevice(dev);
        hipFree(p);
        if (dev != cur) hipSetDevice(cur);
    }
}

}  // namespace

// okano_3_0/Image.cu:14-40
int ImageBasicOp::newImage(Image **outimg)
{
    if (outimg == NULL) return NULL_POINTER;
    ImageCuda *c = new (std::nothrow) ImageCuda;
    if (!c) return OUT_OF_MEM;
    c->imgMeta.width = 0;
    c->imgMeta.height = 0;
    c->imgMeta.roiX1 = c->imgMeta.roiY1 = c->imgMeta.roiX2 = c->imgMeta.roiY2 = 0;
    c->imgMeta.imgData = NULL;
    c->deviceId = -1;
    c->pitchBytes = 0;
    *outimg = &c->imgMeta;
    return NO_ERROR;
}

// okano_3_0/Image.cu:43-91
int ImageBasicOp::deleteImage(Image *inimg)
{
    if (inimg == NULL) return NULL_POINTER;
    ImageCuda *c = IMAGE_CUDA(inimg);
    if (c->deviceId >= device_count()) return OP_OVERFLOW;
    free_data(inimg->imgData, c->deviceId);
    delete c;
    return NO_ERROR;
}

// okano_3_0/Image.cu:94-137
int ImageBasicOp::makeAtCurrentDevice(Image *img, size_t width, size_t height)
{
    if (img == NULL) return NULL_POINTER;
    if (width < 1 || height < 1) return INVALID_DATA;
    if (img->imgData != NULL) return UNMATCH_IMG;
    ImageCuda *c = IMAGE_CUDA(img);
    void *p = NULL;
    size_t pitch = 0;
    if (hipMallocPitch(&p, &pitch, width, height) != hipSuccess) {
        img->imgData = NULL;
        return CUDA_ERROR;
    }
    img->imgData = static_cast<unsigned char *>(p);
    img->width = width;
    img->height = height;
    img->roiX1 = 0;
    img->roiY1 = 0;
    img->roiX2 = (int)width;
    img->roiY2 = (int)height;
    c->deviceId = current_device();
    c->pitchBytes = pitch;
    return NO_ERROR;
}

// okano_3_0/Image.cu:140-175
int ImageBasicOp::makeAtHost(Image *img, size_t width, size_t height)
{
    if (img == NULL) return NULL_POINTER;
    if (width < 1 || height < 1) return INVALID_DATA;
    if (img->imgData != NULL) return UNMATCH_IMG;
    ImageCuda *c = IMAGE_CUDA(img);
    img->imgData = new (std::nothrow) unsigned char[width * height];
    if (img->imgData == NULL) return OUT_OF_MEM;
    img->width = width;
    img->height = height;
    img->roiX1 = 0;
    img->roiY1 = 0;
    img->roiX2 = (int)width;
    img->roiY2 = (int)height;
    c->deviceId = -1;
    c->pitchBytes = width;
    return NO_ERROR;
}

// 8-bit grayscale BMP (okano_3_0/Image.cu:178-290): rows stored bottom-up,
// each padded to 4 bytes; width at 0x12, height at 0x16, bpp at 0x1C, data
// offset at 0x0A.
int ImageBasicOp::readFromFile(const char *filepath, Image *outimg)
{
    if (filepath == NULL || outimg == NULL) return NULL_POINTER;
    ImageCuda *c = IMAGE_CUDA(outimg);
    if (c->deviceId >= device_count()) return OP_OVERFLOW;
    std::ifstream f(filepath, std::ios::in | std::ios::binary);
    if (!f) return NO_FILE;
    char magic[2] = { 0, 0 };
    f.seekg(0x00, std::ios::beg);
    f.read(magic, 2);
    if (!f || magic[0] != 'B' || magic[1] != 'M') return WRONG_FILE;
    uint16_t bpp = 0;
    f.seekg(0x1C, std::ios::beg);
    f.read(reinterpret_cast<char *>(&bpp), 2);
    if (!f || bpp != 8) return WRONG_FILE;
    uint32_t width = 0, height = 0;
    f.seekg(0x12, std::ios::beg);
    f.read(reinterpret_cast<char *>(&width), 4);
    f.read(reinterpret_cast<char *>(&height), 4);
    if (!f || width < 1 || height < 1) return WRONG_FILE;
    uint32_t dataoff = 0;
    f.seekg(0x0A, std::ios::beg);
    f.read(reinterpret_cast<char *>(&dataoff), 4);
    if (!f) return WRONG_FILE;
    unsigned char *data = outimg->imgData;
    bool reuse = true;
    if (outimg->imgData == NULL || c->deviceId >= 0 || outimg->width != width ||
        outimg->height != height) {
        data = new (std::nothrow) unsigned char[(size_t)width * height];
        if (!data) return OUT_OF_MEM;
        reuse = false;
    }
    const uint32_t pad = (4 - (width & 3)) & 3;
    f.seekg(dataoff, std::ios::beg);
    for (uint32_t r = 0; r < height; r++) {
        unsigned char *row = data + (size_t)(height - 1 - r) * width;
        f.read(reinterpret_cast<char *>(row), width);
        if (pad) f.seekg(pad, std::ios::cur);
    }
    if (!f) {
        if (!reuse) delete[] data;
        return WRONG_FILE;
    }
    if (!reuse && outimg->imgData != NULL) free_data(outimg->imgData, c->deviceId);
    outimg->width = width;
    outimg->height = height;
    outimg->roiX1 = 0;
    outimg->roiY1 = 0;
    outimg->roiX2 = (int)width;
    outimg->roiY2 = (int)height;
    outimg->imgData = data;
    c->deviceId = -1;
    c->pitchBytes = width;
    return NO_ERROR;
}

// okano_3_0/Image.cu:293-432.  Like the reference, the whole image is
// written (the header documents ROI-only output, the code writes it all).
int ImageBasicOp::writeToFile(const char *filepath, Image *inimg)
{
    if (filepath == NULL || inimg == NULL) return NULL_POINTER;
    std::ofstream f(filepath, std::ios::out | std::ios::binary);
    if (!f) return NO_FILE;
    int err = copyToHost(inimg);
    if (err < 0) return err;
    ImageCuda *c = IMAGE_CUDA(inimg);
    const uint32_t pad = (4 - (inimg->width & 3)) & 3;
    const uint32_t datalen = (uint32_t)(inimg->height * (inimg->width + pad));
    const uint32_t dataoff = 4 * 256 + 54;
    auto put16 = [&](uint16_t v) { f.write(reinterpret_cast<const char *>(&v), 2); };
    auto put32 = [&](uint32_t v) { f.write(reinterpret_cast<const char *>(&v), 4); };
    put16(0x4D42);                  // "BM"
    put32(datalen + dataoff);       // file size
    put16(0);
    put16(0);
    put32(dataoff);
    put32(40);                      // BITMAPINFOHEADER
    put32((uint32_t)inimg->width);
    put32((uint32_t)inimg->height);
    put16(1);                       // planes
    put16(8);                       // bits per pixel
    put32(0);                       // no compression
    put32(datalen);
    put32(0);
    put32(0);
    put32(256);                     // palette entries
    put32(0);
    for (int i = 0; i < 256; i++) {
        const unsigned char e[4] = { (unsigned char)i, (unsigned char)i, (unsigned char)i, 0xFF };
        f.write(reinterpret_cast<const char *>(e), 4);
    }
    const char zeros[4] = { 0, 0, 0, 0 };
    for (size_t r = 0; r < inimg->height; r++) {
        const unsigned char *row = inimg->imgData + (inimg->height - 1 - r) * c->pitchBytes;
        f.write(reinterpret_cast<const char *>(row), inimg->width);
        if (pad) f.write(zeros, pad);
    }
    return f ? NO_ERROR : NO_FILE;
}

// okano_3_0/Image.cu:435-534
int ImageBasicOp::copyToCurrentDevice(Image *img)
{
    if (img == NULL) return NULL_POINTER;
    ImageCuda *c = IMAGE_CUDA(img);
    if (c->deviceId >= device_count()) return OP_OVERFLOW;
    const int cur = current_device();
    if (img->imgData == NULL || img->width == 0 || img->height == 0 || c->pitchBytes == 0)
        return UNMATCH_IMG;
    if (c->deviceId < 0) {
        void *dev = NULL;
        size_t pitch = 0;
        if (hipMallocPitch(&dev, &pitch, img->width, img->height) != hipSuccess)
            return CUDA_ERROR;
        if (hipMemcpy2D(dev, pitch, img->imgData, c->pitchBytes, img->width, img->height,
                        hipMemcpyHostToDevice) != hipSuccess) {
            hipFree(dev);
            return CUDA_ERROR;
        }
        delete[] img->imgData;
        img->imgData = static_cast<unsigned char *>(dev);
        c->deviceId = cur;
        c->pitchBytes = pitch;
    } else if (c->deviceId != cur) {
        const size_t bytes = c->pitchBytes * img->height;
        void *dev = NULL;
        if (hipMalloc(&dev, bytes) != hipSuccess) return CUDA_ERROR;
        if (hipMemcpyPeer(dev, cur, img->imgData, c->deviceId, bytes) != hipSuccess) {
            hipFree(dev);
            return CUDA_ERROR;
        }
        free_data(img->imgData, c->deviceId);
        img->imgData = static_cast<unsigned char *>(dev);
        c->deviceId = cur;
    }
    return NO_ERROR;
}

// okano_3_0/Image.cu:537-712
int ImageBasicOp::copyToCurrentDevice(Image *srcimg, Image *dstimg)
{
    if (srcimg == NULL) return NULL_POINTER;
    if (dstimg == NULL || dstimg == srcimg) return copyToCurrentDevice(srcimg);
    ImageCuda *s = IMAGE_CUDA(srcimg);
    ImageCuda *d = IMAGE_CUDA(dstimg);
    ImageCuda old = *d;
    bool reuse = true;
    if (srcimg->imgData == NULL || srcimg->width == 0 || srcimg->height == 0 || s->pitchBytes == 0)
        return INVALID_DATA;
    const int ndev = device_count();
    if (s->deviceId >= ndev || d->deviceId >= ndev) return OP_OVERFLOW;
    const int cur = current_device();
    if (d->deviceId != cur) {
        reuse = false;
        dstimg->imgData = NULL;
    } else if (!(((s->deviceId < 0 && srcimg->width == dstimg->width) ||
                  d->pitchBytes == s->pitchBytes) &&
                 srcimg->height == dstimg->height)) {
        reuse = false;
        dstimg->imgData = NULL;
    }
    dstimg->width = srcimg->width;
    dstimg->height = srcimg->height;
    dstimg->roiX1 = srcimg->roiX1;
    dstimg->roiY1 = srcimg->roiY1;
    dstimg->roiX2 = srcimg->roiX2;
    dstimg->roiY2 = srcimg->roiY2;
    d->deviceId = cur;
    if (s->deviceId < 0) {
        if (dstimg->imgData == NULL) {
            void *p = NULL;
            size_t pitch = 0;
            if (hipMallocPitch(&p, &pitch, dstimg->width, dstimg->height) != hipSuccess) {
                *d = old;
                return CUDA_ERROR;
            }
            dstimg->imgData = static_cast<unsigned char *>(p);
            d->pitchBytes = pitch;
        }
        if (hipMemcpy2D(dstimg->imgData, d->pitchBytes, srcimg->imgData, s->pitchBytes,
                        srcimg->width, srcimg->height, hipMemcpyHostToDevice) != hipSuccess) {
            if (!reuse) hipFree(dstimg->imgData);
            *d = old;
            return CUDA_ERROR;
        }
    } else {
        const size_t bytes = s->pitchBytes * srcimg->height;
        if (dstimg->imgData == NULL) {
            void *p = NULL;
            if (hipMalloc(&p, bytes) != hipSuccess) {
                *d = old;
                return CUDA_ERROR;
            }
            dstimg->imgData = static_cast<unsigned char *>(p);
        }
        d->pitchBytes = s->pitchBytes;
        if (hipMemcpyPeer(dstimg->imgData, cur, srcimg->imgData, s->deviceId, bytes) !=
            hipSuccess) {
            if (!reuse) hipFree(dstimg->imgData);
            *d = old;
            return CUDA_ERROR;
        }
    }
    if (old.imgMeta.imgData != NULL && !reuse) free_data(old.imgMeta.imgData, old.deviceId);
    return NO_ERROR;
}

// okano_3_0/Image.cu:715-791
int ImageBasicOp::copyToHost(Image *img)
{
    if (img == NULL) return NULL_POINTER;
    ImageCuda *c = IMAGE_CUDA(img);
    if (c->deviceId >= device_count()) return OP_OVERFLOW;
    const int cur = current_device();
    if (img->imgData == NULL || img->width == 0 || img->height == 0 || c->pitchBytes == 0)
        return UNMATCH_IMG;
    if (c->deviceId < 0) return NO_ERROR;
    unsigned char *host = new (std::nothrow) unsigned char[img->width * img->height];
    if (!host) return OUT_OF_MEM;
    if (c->deviceId != cur) hipSetDevice(c->deviceId);
    hipError_t e = hipMemcpy2D(host, img->width, img->imgData, c->pitchBytes, img->width,
                               img->height, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        delete[] host;
        if (c->deviceId != cur) hipSetDevice(cur);
        return CUDA_ERROR;
    }
    hipFree(img->imgData);
    if (c->deviceId != cur) hipSetDevice(cur);
    img->imgData = host;
    c->deviceId = -1;
    c->pitchBytes = img->width;
    return NO_ERROR;
}

// okano_3_0/Image.cu:794-940
int ImageBasicOp::copyToHost(Image *srcimg, Image *dstimg)
{
    if (srcimg == NULL) return NULL_POINTER;
    if (dstimg == NULL || dstimg == srcimg) return copyToHost(srcimg);
    ImageCuda *s = IMAGE_CUDA(srcimg);
    ImageCuda *d = IMAGE_CUDA(dstimg);
    ImageCuda old = *d;
    bool reuse = true;
    if (srcimg->imgData == NULL || srcimg->width == 0 || srcimg->height == 0 || s->pitchBytes == 0)
        return INVALID_DATA;
    const int ndev = device_count();
    if (s->deviceId >= ndev || d->deviceId >= ndev) return OP_OVERFLOW;
    const int cur = current_device();
    if (d->deviceId >= 0) {
        reuse = false;
        dstimg->imgData = NULL;
    } else if (!(srcimg->width == dstimg->width && srcimg->height == dstimg->height)) {
        reuse = false;
        dstimg->imgData = NULL;
    }
    dstimg->width = srcimg->width;
    dstimg->height = srcimg->height;
    dstimg->roiX1 = srcimg->roiX1;
    dstimg->roiY1 = srcimg->roiY1;
    dstimg->roiX2 = srcimg->roiX2;
    dstimg->roiY2 = srcimg->roiY2;
    d->deviceId = -1;
    d->pitchBytes = dstimg->width;
    if (dstimg->imgData == NULL) {
        dstimg->imgData = new (std::nothrow) unsigned char[srcimg->width * srcimg->height];
        if (dstimg->imgData == NULL) {
            *d = old;
            return OUT_OF_MEM;
        }
    }
    if (s->deviceId < 0) {
        for (size_t r = 0; r < srcimg->height; r++)
            std::memcpy(dstimg->imgData + r * d->pitchBytes, srcimg->imgData + r * s->pitchBytes,
                        srcimg->width);
    } else {
        if (s->deviceId != cur) hipSetDevice(s->deviceId);
        hipError_t e = hipMemcpy2D(dstimg->imgData, d->pitchBytes, srcimg->imgData, s->pitchBytes,
                                   srcimg->width, srcimg->height, hipMemcpyDeviceToHost);
        if (s->deviceId != cur) hipSetDevice(cur);
        if (e != hipSuccess) {
            if (!reuse) delete[] dstimg->imgData;
            *d = old;
            return CUDA_ERROR;
        }
    }
    if (old.imgMeta.imgData != NULL && !reuse) free_data(old.imgMeta.imgData, old.deviceId);
    return NO_ERROR;
}
