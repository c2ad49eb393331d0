// mdas_kdgrid.hpp -- header-only C++ mirror of the reference's
// mdas::KDGrid<Point> (SDK/lib/mdas/kdgrid.h:10-107) and mdas::PointN<D>
// (SDK/lib/mdas/point.h:6-199) over the C ABI of libmdas.so (mdas.h).
//
// The reference apps use KDGrid through these calls only
// (SDK/optixMotionBlur/optixMotionBlur.cpp:331-360, 473-491, 768-806):
//   new KDGrid<Point>(maxSamples, candidatesNum, bitsPerDim, extraImgBits,
//                     alpha, scaleX, scaleY, &log);       // std::ofstream*
//   InitialSampling(); SamplingPass(); Validate();
//   Integrate(float4*, w, h); ConvertToBytes(float4*, uchar4*, w, h);
//   SamplingDensity(float4* host, w, h);
//   reinterpret_cast<float*>(GetSampleCoordinates().Data());
//   float3* values = GetSampleValues().Data();
//   GetNewSamples(); GetNewNodes(); GetNumberOfSamples(); GetNumberOfNodes();
//   GetScaleX(); GetScaleY().
// Those lines compile unchanged against this header with HIP's vector types
// (float2/3/4, uchar4; hip/hip_vector_types.h, included here), and the log
// stream receives the reference's "KEY\nvalue" lines (kdgrid.cu:670-817).
// Failures throw mdas::Error (the reference printed CUDA errors and
// continued).
#pragma once
#include <hip/hip_vector_types.h>

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

static_assert(sizeof(float3) == 12, "a sample value is three packed floats (whitted.h:78)");

// Writes the library's log lines to the app's stream as the reference does
// (`*log << "KEY\n" << value << std::endl`): counts as integers, times as
// floats.
inline void ofstream_log_sink(void* user, const char* key, double value, int is_count) {
    std::ofstream& os = *static_cast<std::ofstream*>(user);
    if (is_count)
        os << key << "\n" << (long long)value << std::endl;
    else
        os << key << "\n" << (float)value << std::endl;
}

template <typename Point>
class KDGrid {
public:
    // kdgrid.h:42-51.  `stream` (an addition) is the HIP stream every call is
    // enqueued on; the reference uses stream 0.
    KDGrid(int maxSamples, int candidatesNum, int bitsPerDim, int extraImgBits, float alpha,
           float scaleX, float scaleY, std::ofstream* log = nullptr, void* stream = nullptr)
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
        c.log_path = nullptr;
        check(mdas_create(&c, &h_));
        if (log) {
            const int rc = mdas_set_log_sink(h_, &ofstream_log_sink, log);
            if (rc < 0) {  // no destructor runs for a constructor that throws
                const std::string m = mdas_last_error();
                mdas_destroy(h_);
                h_ = nullptr;
                throw Error(rc, m);
            }
        }
    }
    ~KDGrid() { mdas_destroy(h_); }
    KDGrid(const KDGrid&) = delete;
    KDGrid& operator=(const KDGrid&) = delete;

    void InitialSampling() { check(mdas_initial_sampling(h_, stream_)); }
    void ComputeErrors() { check(mdas_compute_errors(h_, stream_)); }
    void AdaptiveSampling() { check(mdas_adaptive_sampling(h_, stream_)); }
    void SamplingPass() { check(mdas_sampling_pass(h_, stream_)); }
    // kdgrid.h:59-61: device float4 / uchar4 images, host float4 density
    void Integrate(float4* pixels, int width, int height) {
        check(mdas_integrate(h_, reinterpret_cast<float*>(pixels), width, height, stream_));
    }
    void SamplingDensity(float4* hostPixels, int width, int height) {
        check(mdas_sampling_density(h_, reinterpret_cast<float*>(hostPixels), width, height));
    }
    void ConvertToBytes(float4* pixels, uchar4* pixelsBytes, int width, int height) {
        check(mdas_convert_to_bytes(h_, reinterpret_cast<const float*>(pixels),
                                    reinterpret_cast<unsigned char*>(pixelsBytes), width, height,
                                    stream_));
    }
    bool Validate() {
        int v = 0;
        check(mdas_validate(h_, &v));
        return v != 0;
    }

    // kdgrid.h:65-66: the sample buffers the integrand reads and writes
    DeviceView<Point> GetSampleCoordinates() {
        return {reinterpret_cast<Point*>(mdas_get_sample_coordinates(h_))};
    }
    DeviceView<float3> GetSampleValues() {
        return {reinterpret_cast<float3*>(mdas_get_sample_values(h_))};
    }

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
