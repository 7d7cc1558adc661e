// Image.h -- image object with a region of interest (ROI).
//
// The reference's object model (okano_3_0/Image.h:44-258): an Image is the
// public metadata, an ImageCuda wraps it with the residency (deviceId < 0 =
// host memory, else the HIP device holding it) and the row pitch.  Device
// images are pitched allocations (hipMallocPitch); host images are dense
// (pitch == width).  The ROI is [roiX1, roiX2) x [roiY1, roiY2) -- x2/y2 are
// exclusive, and a newly made image's ROI is the whole image
// (okano_3_0/Image.cu:130).
#ifndef __IMAGE_H__
#define __IMAGE_H__

#include <stddef.h>

#include "ErrorCode.h"

typedef struct Image_st {
    size_t width;            // width in pixels
    size_t height;           // height in pixels
    int roiX1;               // ROI left (inclusive)
    int roiY1;               // ROI top (inclusive)
    int roiX2;               // ROI right (exclusive)
    int roiY2;               // ROI bottom (exclusive)
    unsigned char *imgData;  // pixel data (host or device memory, see ImageCuda)
} Image;

typedef struct ImageCuda_st {
    Image imgMeta;           // the public part
    int deviceId;            // < 0: host memory; else the device holding imgData
    size_t pitchBytes;       // bytes per row (>= width)
} ImageCuda;

// Image* -> its enclosing ImageCuda (every Image comes from newImage).
#define IMAGE_CUDA(img) \
    ((ImageCuda *)((unsigned char *)(img) - offsetof(ImageCuda, imgMeta)))

class ImageBasicOp {
public:
    // Create an empty image (no data, host residency).
    static int newImage(Image **outimg);
    // Free an image and its data.
    static int deleteImage(Image *inimg);
    // Give an empty image a pitched device allocation of width x height on
    // the current device; the ROI becomes the whole image.
    static int makeAtCurrentDevice(Image *img, size_t width, size_t height);
    // Give an empty image a dense host allocation of width x height.
    static int makeAtHost(Image *img, size_t width, size_t height);
    // Read an 8-bit grayscale BMP into host memory (ROI reset to the image).
    static int readFromFile(const char *filepath, Image *outimg);
    // Write the image as an 8-bit grayscale BMP (moves it to host memory).
    static int writeToFile(const char *filepath, Image *inimg);
    // Move the image's data to the current device, in place.
    static int copyToCurrentDevice(Image *img);
    // Copy srcimg into dstimg on the current device.
    static int copyToCurrentDevice(Image *srcimg, Image *dstimg);
    // Move the image's data to host memory, in place.
    static int copyToHost(Image *img);
    // Copy srcimg into dstimg in host memory.
    static int copyToHost(Image *srcimg, Image *dstimg);

    // The ROI as a stand-alone view: width/height are the ROI size, imgData
    // points at the ROI origin, pitch and residency are the image's
    // (okano_3_0/Image.h:218-258).  The view's roi* fields keep the parent's
    // ROI.  Never pass the view to deleteImage.
    static int roiSubImage(Image *inimg, ImageCuda *subimgCud)
    {
        if (inimg == NULL || subimgCud == NULL)
            return NULL_POINTER;
        ImageCuda *inimgCud = IMAGE_CUDA(inimg);
        if (inimg->imgData == NULL || inimg->width == 0 || inimg->height == 0 ||
            inimgCud->pitchBytes == 0)
            return INVALID_DATA;
        subimgCud->imgMeta.width = inimg->roiX2 - inimg->roiX1;
        subimgCud->imgMeta.height = inimg->roiY2 - inimg->roiY1;
        subimgCud->imgMeta.roiX1 = inimg->roiX1;
        subimgCud->imgMeta.roiY1 = inimg->roiY1;
        subimgCud->imgMeta.roiX2 = inimg->roiX2;
        subimgCud->imgMeta.roiY2 = inimg->roiY2;
        subimgCud->imgMeta.imgData =
            inimg->imgData + inimg->roiY1 * inimgCud->pitchBytes + inimg->roiX1;
        subimgCud->deviceId = inimgCud->deviceId;
        subimgCud->pitchBytes = inimgCud->pitchBytes;
        return NO_ERROR;
    }
};

#endif
