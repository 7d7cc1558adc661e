// chain_driver.cpp -- the middle of okano_3_0/SimpleRegionDetect.cu:287-312
// (detectRegion: LocalCluster -> BilateralFilter -> ConnectRegion) written
// the way a caller of the reference's C++ classes writes it: read a BMP, run
// the three stages image to image, write each stage's result as BMP.
//
// usage: chain_driver IN.bmp LC.bmp BF.bmp CR.bmp pnt_range pnt_count gap diffe
//                     sigma_space sigma_range radius min_area max_area
// Prints "status <stage> <code>" per stage; exits 0 when all returned NO_ERROR.
#include <cstdio>
#include <cstdlib>

#include "BilateralFilter.h"
#include "ConnectRegion.h"
#include "Image.h"
#include "LocalCluster.h"

static int write_full(const char *path, Image *img)
{
    img->roiX1 = 0;
    img->roiY1 = 0;
    img->roiX2 = (int)img->width;
    img->roiY2 = (int)img->height;
    return ImageBasicOp::writeToFile(path, img);
}

int main(int argc, char **argv)
{
    if (argc != 14) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    const int pr = std::atoi(argv[5]), pc = std::atoi(argv[6]);
    const int gap = std::atoi(argv[7]), diffe = std::atoi(argv[8]);
    const float ss = (float)std::atof(argv[9]), sr = (float)std::atof(argv[10]);
    const int r = std::atoi(argv[11]);
    const int min_area = std::atoi(argv[12]), max_area = std::atoi(argv[13]);

    Image *in = NULL, *lc = NULL, *bf = NULL, *cr = NULL;
    ImageBasicOp::newImage(&in);
    ImageBasicOp::newImage(&lc);
    ImageBasicOp::newImage(&bf);
    ImageBasicOp::newImage(&cr);
    int rc = ImageBasicOp::readFromFile(argv[1], in);
    std::printf("status read %d\n", rc);
    if (rc == NO_ERROR) {
        LocalCluster cluster((unsigned char)gap, (unsigned char)diffe, pr, pc);
        rc = cluster.localCluster(in, lc);
        std::printf("status localCluster %d\n", rc);
    }
    if (rc == NO_ERROR) {
        BilateralFilter filter(ss, sr, r, 1);
        rc = filter.doFilter(lc, bf);
        std::printf("status doFilter %d\n", rc);
    }
    if (rc == NO_ERROR) {
        ConnectRegion conn(0, max_area, min_area);
        rc = conn.connectRegion(bf, cr);
        std::printf("status connectRegion %d\n", rc);
    }
    if (rc == NO_ERROR) rc = write_full(argv[2], lc);
    if (rc == NO_ERROR) rc = write_full(argv[3], bf);
    if (rc == NO_ERROR) rc = write_full(argv[4], cr);
    ImageBasicOp::deleteImage(in);
    ImageBasicOp::deleteImage(lc);
    ImageBasicOp::deleteImage(bf);
    ImageBasicOp::deleteImage(cr);
    return rc == NO_ERROR ? 0 : 1;
}
