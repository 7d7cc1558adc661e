// LocalCluster.h -- local clustering of gray values (the stage before the
// bilateral filter in okano_3_0/SimpleRegionDetect.cu:287-312).
//
// Same class interface as okano_3_0/LocalCluster.h (members, defaults of
// both constructors, setters and their limits); localCluster runs the
// gfx950 kernel through bf_local_cluster (opencuda_bf.h).  For each pixel
// inside a guard band of pntRange + 2 whose value lies strictly between
// proBlack and proWhite, eight directional walks average the pixels they
// pass while |centre - pixel| <= gapThred and the two side sums differ by at
// most diffeThred; the 8 - pntCount averages farthest from the centre are
// replaced by the centre, and the result is the mean of the pntCount kept.
#ifndef OPENCUDA_AMD_LOCALCLUSTER_H
#define OPENCUDA_AMD_LOCALCLUSTER_H

#include "ErrorCode.h"
#include "Image.h"

class LocalCluster {
protected:
    unsigned char gapThred;     // max gray gap between the centre and a walked pixel
    unsigned char diffeThred;   // max difference of the two side sums
    unsigned char proWhite, proBlack;   // pixels outside (proBlack, proWhite) are copied
    int pntRange;               // points walked per direction (<= 100)
    int pntCount;               // directional averages kept (<= 8)

public:
    LocalCluster()
    {
        this->diffeThred = 40;
        this->gapThred = 40;
        this->pntCount = 7;
        this->pntRange = 16;
        this->proBlack = 10;
        this->proWhite = 250;
    }

    LocalCluster(unsigned char gapthred, unsigned char diffethred, int pntrange, int pntcount,
                 unsigned char problack = 10, unsigned char prowhite = 250)
    {
        this->diffeThred = 10;
        this->gapThred = 10;
        this->pntCount = 5;
        this->pntRange = 8;
        this->proBlack = 10;
        this->proWhite = 250;
        this->setGapThred(gapthred);
        this->setDiffeThred(diffethred);
        this->setPntRange(pntrange);
        this->setPntCount(pntcount);
        this->setProBlack(problack);
        this->setProWhite(prowhite);
    }

    unsigned char getGapThred() const { return this->gapThred; }
    int setGapThred(unsigned char gapthred)
    {
        this->gapThred = gapthred;
        return NO_ERROR;
    }

    unsigned char getDiffeThred() const { return this->diffeThred; }
    int setDiffeThred(unsigned char diffethred)
    {
        this->diffeThred = diffethred;
        return NO_ERROR;
    }

    unsigned char getProBlack() const { return this->proBlack; }
    int setProBlack(unsigned char problack)
    {
        this->proBlack = problack;
        return NO_ERROR;
    }

    unsigned char getProWhite() const { return this->proWhite; }
    int setProWhite(unsigned char prowhite)
    {
        this->proWhite = prowhite;
        return NO_ERROR;
    }

    unsigned char getPntRange() const { return (unsigned char)this->pntRange; }
    int setPntRange(unsigned char pntrange)
    {
        if (pntrange > 100) return INVALID_DATA;
        this->pntRange = pntrange;
        return NO_ERROR;
    }

    unsigned char getPntCount() const { return (unsigned char)this->pntCount; }
    int setPntCount(unsigned char pntcount)
    {
        if (pntcount > 8) return INVALID_DATA;
        this->pntCount = pntcount;
        return NO_ERROR;
    }

    // okano_3_0/LocalCluster.cu:208-274: moves both images to the current
    // device (creating the output with the input ROI's size if it is empty),
    // trims both ROIs to their common size and filters; asynchronous.
    int localCluster(Image *inimg, Image *outimg);
};

#endif
