// host_api_driver.cpp -- host-only paths of the reference-compatible C++
// API (no GPU needed): TemplateFactory pooling, TPLT and BMP round trips,
// BilateralFilter setter validation.  Prints "name value" lines.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "BilateralFilter.h"
#include "Image.h"
#include "Template.h"
#include "TemplateFactory.h"
#include "opencuda_bf.h"

static void out(const char *k, long v) { std::printf("%s %ld\n", k, v); }

// File-format subcommands (the test writes or parses the bytes itself):
//   tplt-read PATH    read PATH, print "rc", "count", then "p x y" per point
//                     and "w <bits>" per weight (IEEE bits in hex)
//   tplt-write PATH   write a fixed 5-point template with negative offsets
//   bmp-read PATH     read PATH, print "rc", "w", "h", the ROI and every row
//                     as hex
//   bmp-write PATH    write a 37 x 11 host image (pixel i = 7 i + 3) whose
//                     ROI is (2, 3)-(30, 9): the whole image is written
static int file_cmd(const char *cmd, const char *path)
{
    if (std::strcmp(cmd, "tplt-read") == 0) {
        Template *t = NULL;
        TemplateBasicOp::newTemplate(&t);
        const int rc = TemplateBasicOp::readFromFile(path, t);
        out("rc", rc);
        if (rc == 0) {
            out("count", (long)t->count);
            for (size_t i = 0; i < t->count; i++)
                std::printf("p %d %d\n", t->tplData[2 * i], t->tplData[2 * i + 1]);
            for (size_t i = 0; i < t->count; i++) {
                uint32_t bits;
                std::memcpy(&bits, &ATTACHED_DATA(t)[i], 4);
                std::printf("w %08x\n", bits);
            }
        }
        TemplateBasicOp::deleteTemplate(t);
        return 0;
    }
    if (std::strcmp(cmd, "tplt-write") == 0) {
        Template *t = NULL;
        TemplateBasicOp::newTemplate(&t);
        TemplateBasicOp::makeAtHost(t, 5);
        const int xy[10] = { 0, 0, -3, 7, 12, -1, -128, 65535, 2147483647, -2147483647 - 1 };
        const float w[5] = { 1.0f, 0.5f, -2.25f, 1e-40f, 3.4e38f };
        std::memcpy(t->tplData, xy, sizeof xy);
        std::memcpy(ATTACHED_DATA(t), w, sizeof w);
        out("rc", TemplateBasicOp::writeToFile(path, t));
        TemplateBasicOp::deleteTemplate(t);
        return 0;
    }
    if (std::strcmp(cmd, "bmp-read") == 0) {
        Image *im = NULL;
        ImageBasicOp::newImage(&im);
        const int rc = ImageBasicOp::readFromFile(path, im);
        out("rc", rc);
        if (rc == 0) {
            out("w", (long)im->width);
            out("h", (long)im->height);
            out("roi", (long)im->roiX1 * 1000000 + im->roiY1 * 10000 + im->roiX2 * 100 + im->roiY2);
            for (size_t r = 0; r < im->height; r++) {
                std::printf("row ");
                for (size_t x = 0; x < im->width; x++) std::printf("%02x", im->imgData[r * im->width + x]);
                std::printf("\n");
            }
        }
        ImageBasicOp::deleteImage(im);
        return 0;
    }
    if (std::strcmp(cmd, "bmp-write") == 0) {
        Image *im = NULL;
        ImageBasicOp::newImage(&im);
        ImageBasicOp::makeAtHost(im, 37, 11);
        for (int i = 0; i < 37 * 11; i++) im->imgData[i] = (unsigned char)(i * 7 + 3);
        im->roiX1 = 2; im->roiY1 = 3; im->roiX2 = 30; im->roiY2 = 9;
        out("rc", ImageBasicOp::writeToFile(path, im));
        ImageBasicOp::deleteImage(im);
        return 0;
    }
    std::fprintf(stderr, "unknown command %s\n", cmd);
    return 2;
}

int main(int argc, char **argv)
{
    if (argc > 2) return file_cmd(argv[1], argv[2]);
    const char *dir = argc > 1 ? argv[1] : ".";
    char path[4096];

    // --- factory: GAUSS template equals the C-ABI generator
    float s1 = 3.0f, s2 = 3.0000005f, s3 = 3.1f, s4 = 2.9999995f;
    Template *a = NULL, *b = NULL, *c = NULL, *d = NULL;
    out("get_gauss", TemplateFactory::getTemplate(&a, TF_SHAPE_GAUSS, 41, &s1));
    std::vector<int32_t> off(41 * 41 * 2);
    std::vector<float> w(41 * 41);
    bf_gauss_template(3.0f, 41, off.data(), w.data());
    out("gauss_count", (long)a->count);
    out("gauss_same_offsets", std::memcmp(a->tplData, off.data(), off.size() * 4) == 0);
    out("gauss_same_weights", std::memcmp(ATTACHED_DATA(a), w.data(), w.size() * 4) == 0);
    // sigma within 1e-6 (same hash set) shares the pooled template
    out("get_gauss_close", TemplateFactory::getTemplate(&b, TF_SHAPE_GAUSS, 41, &s2));
    out("shared_close", a == b);
    out("get_gauss_far", TemplateFactory::getTemplate(&c, TF_SHAPE_GAUSS, 41, &s3));
    out("shared_far", a == c);
    // 2.9999995 is within 1e-6 but hashes to another set: not shared
    out("get_gauss_other_set", TemplateFactory::getTemplate(&d, TF_SHAPE_GAUSS, 41, &s4));
    out("shared_other_set", a == d);
    TemplateFactory::putTemplate(a);
    TemplateFactory::putTemplate(b);
    TemplateFactory::putTemplate(c);
    TemplateFactory::putTemplate(d);

    // --- EUCLIDE
    float sr = 30.0f;
    Template *e = NULL;
    out("get_euclid", TemplateFactory::getTemplate(&e, TF_SHAPE_EUCLIDE, 256, &sr));
    std::vector<float> E(256);
    bf_euclid_template(30.0f, 256, E.data());
    out("euclid_count", (long)e->count);
    out("euclid_same", std::memcmp(ATTACHED_DATA(e), E.data(), 256 * 4) == 0);
    out("euclid_point5", e->tplData[10]);   // point 5 = (5, 0)

    // --- BOX, CIRCLE, ARC: pooled, equal to the C-ABI generators
    {
        Template *bx = NULL, *bx2 = NULL, *ci = NULL, *ci2 = NULL, *ar = NULL;
        out("get_box", TemplateFactory::getTemplate(&bx, TF_SHAPE_BOX, TFSize(5, 3)));
        out("box_count", (long)bx->count);
        std::vector<int32_t> bp(2 * 15);
        std::vector<float> bw(15);
        bf_box_template(5, 3, bp.data(), bw.data());
        out("box_same", std::memcmp(bx->tplData, bp.data(), 15 * 8) == 0 &&
                            std::memcmp(ATTACHED_DATA(bx), bw.data(), 15 * 4) == 0);
        TemplateFactory::getTemplate(&bx2, TF_SHAPE_BOX, TFSize(5, 3));
        out("box_pooled", bx2 == bx);
        out("get_circle", TemplateFactory::getTemplate(&ci, TF_SHAPE_CIRCLE, 9));
        const int nc = bf_circle_template(9, NULL, NULL);
        out("circle_count_match", (long)ci->count == nc);
        std::vector<int32_t> cp(2 * nc);
        std::vector<float> cw(nc);
        bf_circle_template(9, cp.data(), cw.data());
        out("circle_same", std::memcmp(ci->tplData, cp.data(), nc * 8) == 0);
        TemplateFactory::getTemplate(&ci2, TF_SHAPE_CIRCLE, 9);
        out("circle_pooled", ci2 == ci);
        out("get_arc", TemplateFactory::getTemplate(&ar, TF_SHAPE_ARC, TFSize(3, 5)));
        out("arc_count", (long)ar->count);
        Template *bad = NULL;
        out("err_circle_even", TemplateFactory::getTemplate(&bad, TF_SHAPE_CIRCLE, 8));
        out("err_arc_order", TemplateFactory::getTemplate(&bad, TF_SHAPE_ARC, TFSize(5, 3)));
        TemplateFactory::putTemplate(bx);
        TemplateFactory::putTemplate(bx2);
        TemplateFactory::putTemplate(ci);
        TemplateFactory::putTemplate(ci2);
        TemplateFactory::putTemplate(ar);
    }

    // --- errors
    Template *z = NULL;
    float zero = 0.0f;
    out("err_shape", TemplateFactory::getTemplate(&z, 7, 41, &s1));
    out("err_box", TemplateFactory::getTemplate(&z, TF_SHAPE_BOX, 41, &s1));
    out("err_even", TemplateFactory::getTemplate(&z, TF_SHAPE_GAUSS, 40, &s1));
    out("err_sigma0", TemplateFactory::getTemplate(&z, TF_SHAPE_GAUSS, 41, &zero));
    out("err_null", TemplateFactory::getTemplate(NULL, TF_SHAPE_GAUSS, 41, &s1));
    out("err_put_null", TemplateFactory::putTemplate(NULL));

    // --- TPLT round trip
    std::snprintf(path, sizeof path, "%s/gauss.tplt", dir);
    Template *g = NULL;
    TemplateFactory::getTemplate(&g, TF_SHAPE_GAUSS, 41, &s1);
    out("tplt_write", TemplateBasicOp::writeToFile(path, g));
    Template *r = NULL;
    TemplateBasicOp::newTemplate(&r);
    out("tplt_read", TemplateBasicOp::readFromFile(path, r));
    out("tplt_count", (long)r->count);
    out("tplt_same_points", std::memcmp(r->tplData, g->tplData, g->count * 8) == 0);
    out("tplt_same_weights", std::memcmp(ATTACHED_DATA(r), ATTACHED_DATA(g), g->count * 4) == 0);
    out("tplt_missing", TemplateBasicOp::readFromFile("/nonexistent/x.tplt", r));
    TemplateBasicOp::deleteTemplate(r);
    TemplateFactory::putTemplate(g);

    // --- BMP round trip of a host image with odd width
    Image *im = NULL, *back = NULL;
    ImageBasicOp::newImage(&im);
    ImageBasicOp::newImage(&back);
    out("make_host", ImageBasicOp::makeAtHost(im, 37, 11));
    for (int i = 0; i < 37 * 11; i++) im->imgData[i] = (unsigned char)(i * 7 + 3);
    out("roi_default", im->roiX2 == 37 && im->roiY2 == 11);
    std::snprintf(path, sizeof path, "%s/img.bmp", dir);
    out("bmp_write", ImageBasicOp::writeToFile(path, im));
    out("bmp_read", ImageBasicOp::readFromFile(path, back));
    out("bmp_same", back->width == 37 && back->height == 11 &&
                        std::memcmp(back->imgData, im->imgData, 37 * 11) == 0);
    out("bmp_missing", ImageBasicOp::readFromFile("/nonexistent/x.bmp", back));
    out("make_twice", ImageBasicOp::makeAtHost(im, 5, 5));
    ImageCuda sub = {};
    im->roiX1 = 2; im->roiY1 = 3; im->roiX2 = 30; im->roiY2 = 9;
    out("roi_sub", ImageBasicOp::roiSubImage(im, &sub));
    out("roi_sub_w", (long)sub.imgMeta.width);
    out("roi_sub_h", (long)sub.imgMeta.height);
    out("roi_sub_off", (long)(sub.imgMeta.imgData - im->imgData));
    ImageBasicOp::deleteImage(im);
    ImageBasicOp::deleteImage(back);

    // --- allocation failure: a 2^60-byte host image cannot be allocated;
    // the image stays empty and the next allocation works
    Image *huge = NULL;
    ImageBasicOp::newImage(&huge);
    out("make_host_huge", ImageBasicOp::makeAtHost(huge, (size_t)1 << 40, (size_t)1 << 20));
    out("huge_still_empty", huge->imgData == NULL);
    out("make_host_after", ImageBasicOp::makeAtHost(huge, 16, 16));
    ImageBasicOp::deleteImage(huge);

    // --- BilateralFilter setters (okano_3_0/BilateralFilter.h:121-228)
    BilateralFilter bf;
    out("default_radius", bf.getRadius());
    out("set_radius0", bf.setRadius(0));
    out("set_radius21", bf.setRadius(21));
    out("set_radius20", bf.setRadius(20));
    out("set_repeat0", bf.setRepeat(0));
    out("set_sigma_neg", bf.setSigmaSpace(-1.0f));
    out("sigma_kept", bf.getSigmaSpace() == 0.0f);
    out("set_sigma", bf.setSigmaSpace(5.0f));
    out("set_range", bf.setSigmaRange(25.0f));
    out("dofilter_null", bf.doFilter(NULL, NULL));
    out("dofilter_repeat0", BilateralFilter(5.0f, 25.0f, 3, 0).doFilter(NULL));
    BilateralFilter bad(-1.0f, 25.0f, 3, 1);
    Image *dummy = NULL;
    ImageBasicOp::newImage(&dummy);
    out("dofilter_no_template", bad.doFilter(dummy, dummy));
    ImageBasicOp::deleteImage(dummy);
    return 0;
}
