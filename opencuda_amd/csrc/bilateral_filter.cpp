// bilateral_filter.cpp -- BilateralFilter::doFilter over the C ABI.
//
// Staging follows okano_3_0/BilateralFilter.cu:202-345: the input (and
// output) move to the current device in place, the output is created with
// the input ROI's size when it is empty, both ROIs are trimmed to their
// common size, and the filter runs on the null stream without a host wait.
// The spatial and range templates become one device table set (bf_lut),
// cached per (template pair, device) until either template is deleted.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "BilateralFilter.h"
#include "Template.h"
#include "opencuda_bf.h"

namespace bfimpl {

namespace {

struct CacheKey {
    const Template *g;
    const Template *e;
    int device;
    bool operator<(const CacheKey &o) const
    {
        return std::tie(g, e, device) < std::tie(o.g, o.e, o.device);
    }
};

std::mutex &cache_mu()
{
    static std::mutex *m = new std::mutex();
    return *m;
}

std::map<CacheKey, bf_lut> &cache()
{
    static auto *c = new std::map<CacheKey, bf_lut>();
    return *c;
}

// Host copies of a template's points and weights (from either residency).
int read_template(Template *t, std::vector<int32_t> &xy, std::vector<float> &w)
{
    TemplateCuda *c = TEMPLATE_CUDA(t);
    if (t->tplData == NULL || t->count == 0) return INVALID_DATA;
    xy.resize(2 * t->count);
    w.resize(t->count);
    if (c->deviceId < 0) {
        std::copy(t->tplData, t->tplData + 2 * t->count, xy.begin());
        std::copy(c->attachedData, c->attachedData + t->count, w.begin());
        return NO_ERROR;
    }
    if (hipMemcpy(xy.data(), t->tplData, 2 * t->count * sizeof(int), hipMemcpyDeviceToHost) !=
            hipSuccess ||
        hipMemcpy(w.data(), c->attachedData, t->count * sizeof(float), hipMemcpyDeviceToHost) !=
            hipSuccess)
        return CUDA_ERROR;
    return NO_ERROR;
}

int get_lut(Template *g, Template *e, bf_lut *out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CUDA_ERROR;
    std::lock_guard<std::mutex> lock(cache_mu());
    CacheKey key{ g, e, dev };
    auto it = cache().find(key);
    if (it != cache().end()) {
        *out = it->second;
        return NO_ERROR;
    }
    std::vector<int32_t> gxy, exy;
    std::vector<float> gw, ew;
    int rc = read_template(g, gxy, gw);
    if (rc == NO_ERROR) rc = read_template(e, exy, ew);
    if (rc != NO_ERROR) return rc;
    if (ew.size() < 256) return INVALID_DATA;
    bf_lut lut = NULL;
    rc = bf_lut_create(gxy.data(), gw.data(), (int32_t)gw.size(), ew.data(), &lut);
    if (rc != NO_ERROR) return rc;
    cache().emplace(key, lut);
    *out = lut;
    return NO_ERROR;
}

}  // namespace

void lut_cache_forget(const Template *tpl)
{
    std::lock_guard<std::mutex> lock(cache_mu());
    for (auto it = cache().begin(); it != cache().end();) {
        if (it->first.g == tpl || it->first.e == tpl) {
            bf_lut_destroy(it->second);
            it = cache().erase(it);
        } else {
            ++it;
        }
    }
}

static bf_image_view view_of(const ImageCuda &sub)
{
    bf_image_view v;
    v.data = sub.imgMeta.imgData;
    v.width = (int32_t)sub.imgMeta.width;
    v.height = (int32_t)sub.imgMeta.height;
    v.pitch = (int64_t)sub.pitchBytes;
    return v;
}

}  // namespace bfimpl

using namespace bfimpl;

// okano_3_0/BilateralFilter.cu:202-254, with the racy in-place launches
// replaced by `repeat` snapshot passes.
int BilateralFilter::doFilter(Image *inoutimg)
{
    if (repeat <= 0) return INVALID_DATA;
    if (inoutimg == NULL) return NULL_POINTER;
    if (gaussian == NULL || euclid == NULL) return INVALID_DATA;
    ImageBasicOp::copyToCurrentDevice(inoutimg);   // result checked via roiSubImage
    bf_lut lut = NULL;
    int err = get_lut(gaussian, euclid, &lut);
    if (err != NO_ERROR) return err;
    err = TemplateBasicOp::copyToCurrentDevice(gaussian);
    if (err != NO_ERROR) return err;
    err = TemplateBasicOp::copyToCurrentDevice(euclid);
    if (err != NO_ERROR) return err;
    ImageCuda sub;
    err = ImageBasicOp::roiSubImage(inoutimg, &sub);
    if (err != NO_ERROR) return err;
    if (radius <= 0 || radius > DEF_FILTER_RANGE) return NO_ERROR;   // kernel exits (:75-76)
    bf_image_view v = view_of(sub);
    if (v.width <= 0 || v.height <= 0) return NO_ERROR;
    return bf_exact_inplace(&v, lut, radius, mode, repeat, NULL, NULL);
}

// okano_3_0/BilateralFilter.cu:257-345
int BilateralFilter::doFilter(Image *inimg, Image *outimg)
{
    if (repeat <= 0) return INVALID_DATA;
    if (inimg == NULL || outimg == NULL) return NULL_POINTER;
    if (gaussian == NULL || euclid == NULL) return INVALID_DATA;
    ImageBasicOp::copyToCurrentDevice(inimg);      // the reference ignores this result
    bf_lut lut = NULL;
    int err = get_lut(gaussian, euclid, &lut);
    if (err != NO_ERROR) return err;
    err = TemplateBasicOp::copyToCurrentDevice(gaussian);
    if (err != NO_ERROR) return err;
    err = TemplateBasicOp::copyToCurrentDevice(euclid);
    if (err != NO_ERROR) return err;
    ImageCuda insub;
    err = ImageBasicOp::roiSubImage(inimg, &insub);
    if (err != NO_ERROR) return err;
    err = ImageBasicOp::copyToCurrentDevice(outimg);
    if (err != NO_ERROR) {
        err = ImageBasicOp::makeAtCurrentDevice(outimg, inimg->roiX2 - inimg->roiX1,
                                                inimg->roiY2 - inimg->roiY1);
        if (err != NO_ERROR) return err;
    }
    ImageCuda outsub;
    err = ImageBasicOp::roiSubImage(outimg, &outsub);
    if (err != NO_ERROR) return err;
    if (radius <= 0 || radius > DEF_FILTER_RANGE) return NO_ERROR;   // kernel exits (:75-76)
    bf_image_view vin = view_of(insub), vout = view_of(outsub);
    if (vin.width <= 0 || vin.height <= 0 || vout.width <= 0 || vout.height <= 0)
        return NO_ERROR;
    // Output size = min(in, out) is applied by bf_exact; repeat > 1 relaunches
    // the identical kernel (:336-341), which changes nothing.
    return bf_exact(&vin, &vout, lut, radius, mode, NULL);
}
