// The reference's C++ smoke test shape (tdctest/tdctest.cpp:21-59) against the
// drop-in class: a 20 x 20 grid of zeros, 750 chirps of 1500 zero samples, a
// 5 Hz Doppler band around the centroid with the range window on, then
// start(). Exit status 0 when imageBuffer() is non-null after start() (the
// reference's assertion) and, the inputs being all zero, the image is all
// zero; 1 otherwise. Prints "smoke_cpp ok" on success.
#include <complex>
#include <cstdio>
#include <vector>

#include "tdcprocessor.h"

int main()
{
    TdcProcessor tdc(0);
    const int nRows = 20, nCols = 20;
    std::vector<float> grid(static_cast<size_t>(nRows) * nCols * 3);
    tdc.setFocusGrid(grid.data(), nRows, nCols);

    const int nPri = 750, nSamples = 1500;
    const size_t n = static_cast<size_t>(nPri) * nSamples;
    std::vector<std::complex<float>> raw(n);
    std::vector<float> priTimes(nPri), sampleTimes(nSamples), position(n * 3), velocity(n * 3),
        attitude(n * 4);
    float3 bore;
    bore.x = 0.0F;
    bore.y = 1.0F;
    bore.z = 0.0F;
    tdc.setRawData(raw.data(), priTimes.data(), sampleTimes.data(), position.data(),
                   velocity.data(), attitude.data(), nPri, nSamples, 1.0F, 1.0F, bore);
    tdc.setLoggerSinks({std::make_shared<tdclog::stdout_sink>()});
    tdc.start(0.0F, 5.0F, true, true);

    const float2 *img = tdc.imageBuffer();
    if (img == nullptr) {
        std::printf("smoke_cpp: imageBuffer() is null after start()\n");
        return 1;
    }
    for (int i = 0; i < nRows * nCols; ++i) {
        if (img[i].x != 0.0F || img[i].y != 0.0F) {
            std::printf("smoke_cpp: pixel %d is (%g, %g), expected zero\n", i, img[i].x, img[i].y);
            return 1;
        }
    }
    std::printf("smoke_cpp ok\n");
    return 0;
}
