// region_api.cpp -- LocalCluster::localCluster and
// ConnectRegion::connectRegion over the C ABI.
//
// Staging follows okano_3_0/LocalCluster.cu:208-274 and
// okano_3_0/ConnectRegion.cu:789-854: the input (and output) move to the
// current device in place, the output is created with the input ROI's size
// when it is empty, both ROIs are trimmed to their common size, and the
// kernels run on the null stream without a host wait.
#include "ConnectRegion.h"
#include "LocalCluster.h"
#include "opencuda_bf.h"

namespace {

bf_image_view view_of(const ImageCuda &sub)
{
    bf_image_view v;
    v.data = sub.imgMeta.imgData;
    v.width = (int32_t)sub.imgMeta.width;
    v.height = (int32_t)sub.imgMeta.height;
    v.pitch = (int64_t)sub.pitchBytes;
    return v;
}

// Device residency for both images and their ROI views (the reference's
// common preamble).
int stage(Image *inimg, Image *outimg, ImageCuda *insub, ImageCuda *outsub)
{
    int err = ImageBasicOp::copyToCurrentDevice(inimg);
    if (err != NO_ERROR) return err;
    err = ImageBasicOp::copyToCurrentDevice(outimg);
    if (err != NO_ERROR) {
        err = ImageBasicOp::makeAtCurrentDevice(outimg, inimg->roiX2 - inimg->roiX1,
                                                inimg->roiY2 - inimg->roiY1);
        if (err != NO_ERROR) return err;
    }
    err = ImageBasicOp::roiSubImage(inimg, insub);
    if (err != NO_ERROR) return err;
    return ImageBasicOp::roiSubImage(outimg, outsub);
}

}  // namespace

int LocalCluster::localCluster(Image *inimg, Image *outimg)
{
    if (inimg == NULL || inimg->imgData == NULL) return NULL_POINTER;
    if (outimg == NULL) return NULL_POINTER;
    ImageCuda insub, outsub;
    int err = stage(inimg, outimg, &insub, &outsub);
    if (err != NO_ERROR) return err;
    bf_image_view vin = view_of(insub), vout = view_of(outsub);
    if (vin.width <= 0 || vin.height <= 0 || vout.width <= 0 || vout.height <= 0) return NO_ERROR;
    // the allocation holding the ROI: pitch x full height from the image origin
    const ImageCuda *full = IMAGE_CUDA(inimg);
    const int64_t origin = (int64_t)inimg->roiY1 * (int64_t)full->pitchBytes + inimg->roiX1;
    const int64_t lo = -origin, hi = (int64_t)full->pitchBytes * (int64_t)inimg->height - origin;
    bf_local_cluster_params p;
    p.pnt_range = pntRange;
    p.pnt_count = pntCount;
    p.gap_thred = gapThred;
    p.diffe_thred = diffeThred;
    p.pro_black = proBlack;
    p.pro_white = proWhite;
    return bf_local_cluster(&vin, lo, hi, &vout, &p, NULL);
}

int ConnectRegion::connectRegion(Image *inimg, Image *outimg)
{
    if (inimg == NULL || outimg == NULL) return NULL_POINTER;
    ImageCuda insub, outsub;
    int err = stage(inimg, outimg, &insub, &outsub);
    if (err != NO_ERROR) return err;
    bf_image_view vin = view_of(insub), vout = view_of(outsub);
    if (vin.width <= 0 || vin.height <= 0 || vout.width <= 0 || vout.height <= 0) return NO_ERROR;
    return bf_connect_region(&vin, &vout, minArea, maxArea, 10, 250, NULL);
}
