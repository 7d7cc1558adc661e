// ConnectRegion.h -- connected-region labelling (the stage after the
// bilateral filter in okano_3_0/SimpleRegionDetect.cu:287-312).
//
// Same class interface as okano_3_0/ConnectRegion.h; connectRegion runs the
// gfx950 kernels through bf_connect_region (opencuda_bf.h) with the gray
// range the reference fixes (10..250).  Pixels in that range form
// 8-connected regions; regions with an area outside [minArea, maxArea] are
// dropped; the others are numbered in raster order of their first pixel and
// the output pixel is number + 1 as u8 (0 = no region).  `threshold` is
// kept for interface compatibility: the reference never uses it either.
#ifndef OPENCUDA_AMD_CONNECTREGION_H
#define OPENCUDA_AMD_CONNECTREGION_H

#include "ErrorCode.h"
#include "Image.h"

class ConnectRegion {
protected:
    int threshold;
    int maxArea, minArea;

public:
    ConnectRegion()
    {
        this->threshold = 0;
        this->maxArea = 100000;
        this->minArea = 100;
    }

    ConnectRegion(int threshold, int maxArea, int minArea)
    {
        this->threshold = 0;
        this->maxArea = 100000;
        this->minArea = 100;
        setThreshold(threshold);
        setMaxArea(maxArea);
        setMinArea(minArea);
    }

    int getThreshold() const { return this->threshold; }
    int setThreshold(int threshold)
    {
        this->threshold = threshold;
        return NO_ERROR;
    }

    int getMinArea() const { return this->minArea; }
    int setMinArea(int minArea)
    {
        this->minArea = minArea;
        return NO_ERROR;
    }

    int getMaxArea() const { return this->maxArea; }
    int setMaxArea(int maxArea)
    {
        this->maxArea = maxArea;
        return NO_ERROR;
    }

    // okano_3_0/ConnectRegion.cu:789-1043: moves both images to the current
    // device (creating the output with the input ROI's size if it is empty),
    // trims both ROIs to their common size and labels; asynchronous.
    int connectRegion(Image *inimg, Image *outimg);
};

#endif
