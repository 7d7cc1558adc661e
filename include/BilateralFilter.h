// BilateralFilter.h -- exact bilateral filter on MI355X.
//
// Drop-in for okano_3_0/BilateralFilter.h:32-247: same members, setters,
// defaults and doFilter entry points.  Setters validate as the reference
// does (sigma > 0, 1 <= radius <= DEF_FILTER_RANGE, repeat >= 1) and leave
// the old value on error; the 4-argument constructor ignores setter errors.
// The spatial template is always the 41x41 ring (radius 20) and the range
// table has 256 entries, both from TemplateFactory.
//
// doFilter(in, out) filters min(in ROI, out ROI) pixels; the output is
// created on the current device with the input ROI's size if it is empty.
// Both images end up resident on the current device.  It returns before the
// filter finishes (stream order on the null stream); copyToHost waits.
// doFilter(inout) filters in place `repeat` times, each pass reading a
// snapshot of the previous one.
//
// Extension (not in the reference): setMode(BF_MODE_STD) switches to the
// own-window filter; the default BF_MODE_REF reproduces the reference kernel.
#ifndef __BILATERALFILTER_H__
#define __BILATERALFILTER_H__

#include "ErrorCode.h"
#include "Image.h"
#include "TemplateFactory.h"

#define DEF_FILTER_RANGE 20

class BilateralFilter {
protected:
    Template *gaussian, *euclid;   // spatial template (41x41), range table (256)
    float sigmaSpace, sigmaRange;
    int radius;
    int repeat;
    int mode;                      // BF_MODE_REF (0) or BF_MODE_STD (1)

public:
    BilateralFilter()
        : gaussian(NULL), euclid(NULL), sigmaSpace(0.0f), sigmaRange(0.0f),
          radius(0), repeat(0), mode(0) {}

    BilateralFilter(float sigmaspace, float sigmarange, int radius, int repeat)
        : gaussian(NULL), euclid(NULL), sigmaSpace(0.0f), sigmaRange(0.0f),
          radius(0), repeat(0), mode(0)
    {
        this->setSigmaSpace(sigmaspace);
        this->setSigmaRange(sigmarange);
        this->setRadius(radius);
        this->setRepeat(repeat);
    }

    ~BilateralFilter()
    {
        if (gaussian != NULL)
            TemplateFactory::putTemplate(gaussian);
        if (euclid != NULL)
            TemplateFactory::putTemplate(euclid);
    }

    float getSigmaSpace() const { return sigmaSpace; }

    int setSigmaSpace(float sigmaspace)
    {
        if (sigmaspace <= 0)
            return INVALID_DATA;
        this->sigmaSpace = sigmaspace;
        if (gaussian != NULL)
            TemplateFactory::putTemplate(gaussian);
        gaussian = NULL;
        return TemplateFactory::getTemplate(&gaussian, TF_SHAPE_GAUSS,
                                            2 * DEF_FILTER_RANGE + 1, &sigmaSpace);
    }

    float getSigmaRange() const { return sigmaRange; }

    int setSigmaRange(float sigmarange)
    {
        if (sigmarange <= 0)
            return INVALID_DATA;
        this->sigmaRange = sigmarange;
        if (euclid != NULL)
            TemplateFactory::putTemplate(euclid);
        euclid = NULL;
        return TemplateFactory::getTemplate(&euclid, TF_SHAPE_EUCLIDE, 256, &sigmaRange);
    }

    int getRadius() const { return radius; }

    int setRadius(int radius)
    {
        if (radius <= 0 || radius > DEF_FILTER_RANGE)
            return INVALID_DATA;
        this->radius = radius;
        return NO_ERROR;
    }

    int getRepeat() const { return repeat; }

    int setRepeat(int repeat)
    {
        if (repeat <= 0)
            return INVALID_DATA;
        this->repeat = repeat;
        return NO_ERROR;
    }

    int getMode() const { return mode; }

    int setMode(int mode)
    {
        if (mode != 0 && mode != 1)
            return INVALID_DATA;
        this->mode = mode;
        return NO_ERROR;
    }

    // In place, `repeat` passes (okano_3_0/BilateralFilter.cu:202-254).
    int doFilter(Image *inoutimg);
    // Out of place (okano_3_0/BilateralFilter.cu:257-345).
    int doFilter(Image *inimg, Image *outimg);
};

#endif
