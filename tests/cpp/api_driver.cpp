// api_driver.cpp -- exercises the reference-compatible C++ API the way a
// caller of okano_3_0/BilateralFilter.h would: read a BMP, set a ROI, run
// BilateralFilter::doFilter, write the result as BMP.
//
// usage: api_driver IN.bmp OUT.bmp sigma_space sigma_range radius repeat
//                   x1 y1 x2 y2 out_w out_h mode inplace
//   out_w/out_h > 0: pre-create the output on the device with that size
//   (its ROI = whole image), else let doFilter create it.
//   inplace = 1: doFilter(in) and write `in`.
// Prints "status <code>" and exits 0 when the call returned NO_ERROR.
//
// usage: api_driver alloc-fail
//   asks for a 1 TiB device image (more than the device holds): prints
//   "device_huge <code>", then "device_after <code>" for a small image
//   allocated next.  No kernel runs.
#include <cstdio>
#include <cstdlib>

#include "BilateralFilter.h"
#include "Image.h"

#include <cstring>

static int alloc_fail()
{
    Image *img = NULL;
    ImageBasicOp::newImage(&img);
    std::printf("device_huge %d\n",
                ImageBasicOp::makeAtCurrentDevice(img, (size_t)1 << 20, (size_t)1 << 20));
    std::printf("device_empty %d\n", img->imgData == NULL ? 1 : 0);
    std::printf("device_after %d\n", ImageBasicOp::makeAtCurrentDevice(img, 64, 64));
    ImageBasicOp::deleteImage(img);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 2 && std::strcmp(argv[1], "alloc-fail") == 0) return alloc_fail();
    if (argc != 15) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    const float ss = (float)std::atof(argv[3]), sr = (float)std::atof(argv[4]);
    const int r = std::atoi(argv[5]), repeat = std::atoi(argv[6]);
    const int x1 = std::atoi(argv[7]), y1 = std::atoi(argv[8]);
    const int x2 = std::atoi(argv[9]), y2 = std::atoi(argv[10]);
    const int ow = std::atoi(argv[11]), oh = std::atoi(argv[12]);
    const int mode = std::atoi(argv[13]), inplace = std::atoi(argv[14]);

    Image *in = NULL, *out = NULL;
    ImageBasicOp::newImage(&in);
    ImageBasicOp::newImage(&out);
    int rc = ImageBasicOp::readFromFile(argv[1], in);
    if (rc != NO_ERROR) {
        std::printf("status %d\n", rc);
        return 1;
    }
    in->roiX1 = x1;
    in->roiY1 = y1;
    in->roiX2 = x2;
    in->roiY2 = y2;
    if (ow > 0 && oh > 0) ImageBasicOp::makeAtCurrentDevice(out, ow, oh);

    BilateralFilter bf(ss, sr, r, repeat);
    bf.setMode(mode);
    rc = inplace ? bf.doFilter(in) : bf.doFilter(in, out);
    std::printf("status %d\n", rc);
    if (rc == NO_ERROR) {
        Image *res = inplace ? in : out;
        res->roiX1 = 0;
        res->roiY1 = 0;
        res->roiX2 = (int)res->width;
        res->roiY2 = (int)res->height;
        int wrc = ImageBasicOp::writeToFile(argv[2], res);
        if (wrc != NO_ERROR) {
            std::printf("write %d\n", wrc);
            rc = wrc;
        }
    }
    ImageBasicOp::deleteImage(in);
    ImageBasicOp::deleteImage(out);
    return rc == NO_ERROR ? 0 : 1;
}
