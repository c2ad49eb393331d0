// mdas_kdgrid.hpp -- header-only C++ mirror of the reference's
// mdas::KDGrid<Point> (SDK/lib/mdas/kdgrid.h:10-107) and mdas::PointN<D>
// (SDK/lib/mdas/point.h:6-199) over the C ABI of libmdas.so (mdas.h).
//
// The reference apps use KDGrid through these calls only
// (SDK/optixMotionBlur/optixMotionBlur.cpp:331-360, 473-491, 768-806):
//   new KDGrid<Point>(maxSamples, candidatesNum, bitsPerDim, extraImgBits,
//                     alpha, scaleX, scaleY, &log);
//   InitialSampling(); SamplingPass(); Integrate(px, w, h);
//   ConvertToBytes(px, bytes, w, h); SamplingDensity(hostPx, w, h);
//   Validate(); GetSampleCoordinates().Data(); GetSampleValues().Data();
//   GetNewSamples(); GetNumberOfSamples(); GetNumberOfNodes(); GetScaleX/Y().
// Those compile unchanged against this header.  Failures throw
// mdas::Error (the reference printed CUDA errors and continued).
#pragma once
#include <fstream>
#include <stdexcept>
#include <string>

#include "mdas.h"

namespace mdas {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc < 0) throw Error(rc, mdas_last_error());
}

// Storage layout of one sample point: `D` floats (point.h:6-11).
template <int D>
struct PointN {
    static const int DIM = D;
    float data[D];
    float& operator[](int i) { return data[i]; }
    const float& operator[](int i) const { return data[i]; }
};
typedef PointN<3> Point3;
typedef PointN<4> Point4;
typedef PointN<5> Point5;
typedef PointN<6> Point6;

// Non-owning view of a device buffer (the reference's Buffer<T>::Data(),
// buffer.h:63-65).
template <typename T>
struct DeviceView {
    T* ptr;
    T* Data() { return ptr; }
};

// float3 layout of a sample value (SDK/cuda/whitted.h:80).
struct Value3 {
    float x, y, z;
};

template <typename Point>
class KDGrid {
public:
    // kdgrid.h:42-51.  The log stream is replaced by a log file path: the
    // library appends the reference's "KEY\nvalue" lines to it.
    KDGrid(int maxSamples, int candidatesNum, int bitsPerDim, int extraImgBits, float alpha,
           float scaleX, float scaleY, const char* logPath = nullptr, void* stream = nullptr)
        : stream_(stream) {
        mdas_config c;
        c.dim = Point::DIM;
        c.max_samples = maxSamples;
        c.candidates_num = candidatesNum;
        c.bits_per_dim = bitsPerDim;
        c.extra_img_bits = extraImgBits;
        c.alpha = alpha;
        c.scale_x = scaleX;
        c.scale_y = scaleY;
        c.log_path = logPath;
        check(mdas_create(&c, &h_));
    }
    ~KDGrid() { mdas_destroy(h_); }
    KDGrid(const KDGrid&) = delete;
    KDGrid& operator=(const KDGrid&) = delete;

    void InitialSampling() { check(mdas_initial_sampling(h_, stream_)); }
    void ComputeErrors() { check(mdas_compute_errors(h_, stream_)); }
    void AdaptiveSampling() { check(mdas_adaptive_sampling(h_, stream_)); }
    void SamplingPass() { check(mdas_sampling_pass(h_, stream_)); }
    // float4 device buffers (CUDAOutputBuffer<float4> in the reference apps)
    void Integrate(void* pixels, int width, int height) {
        check(mdas_integrate(h_, (float*)pixels, width, height, stream_));
    }
    void SamplingDensity(void* hostPixels, int width, int height) {
        check(mdas_sampling_density(h_, (float*)hostPixels, width, height));
    }
    void ConvertToBytes(void* pixels, void* pixelsBytes, int width, int height) {
        check(mdas_convert_to_bytes(h_, (const float*)pixels, (unsigned char*)pixelsBytes, width,
                                    height, stream_));
    }
    bool Validate() {
        int v = 0;
        check(mdas_validate(h_, &v));
        return v != 0;
    }

    DeviceView<Point> GetSampleCoordinates() { return {(Point*)mdas_get_sample_coordinates(h_)}; }
    DeviceView<Value3> GetSampleValues() { return {(Value3*)mdas_get_sample_values(h_)}; }

    int GetCandidatesNum() { return mdas_get_candidates_num(h_); }
    int GetMaxNodeSize() { return mdas_get_max_node_size(h_); }
    int GetNewSamples() { return mdas_get_new_samples(h_); }
    int GetNewNodes() { return mdas_get_new_nodes(h_); }
    int GetNumberOfSamples() { return mdas_get_number_of_samples(h_); }
    int GetNumberOfNodes() { return mdas_get_number_of_nodes(h_); }
    int GetMaxSamples() { return mdas_get_max_samples(h_); }
    float GetScaleX() { return mdas_get_scale_x(h_); }
    float GetScaleY() { return mdas_get_scale_y(h_); }

    // additions: the handle for the rest of the C ABI, the stream in use
    mdas_handle handle() { return h_; }
    void* stream() { return stream_; }

private:
    mdas_handle h_ = nullptr;
    void* stream_ = nullptr;
};

}  // namespace mdas
