// DeviceInterface.h -- host-side device interface of the plug-in layer.
//
// Mirrors BehaveRT::DeviceInterface (Core/include/DeviceInterface.cuh:53-297): memory
// management, a thread-sync switch and kernelCall for custom plug-ins.  The hot
// path does not go through kernelCall: the plug-ins call the C ABI
// (include/behavert_amd.h), which owns the gfx950 kernels.  Errors throw
// BehaveRT::DeviceError instead of exiting the process (CUDA_SAFE_CALL,
// Core/DeviceInterface.cu:106).
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/behavert_amd.h"

namespace BehaveRT {

class DeviceError : public std::runtime_error {
public:
    explicit DeviceError(const std::string& m) : std::runtime_error(m) {}
};

inline void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

/// Simulation parameters of the device interface (SimParams in the reference).
struct SimParams {
    int commonBlockDim = 256;
    bool useThreadSync = false;
};

typedef void (*genericKernelFuncPointer)();
typedef void (*kernelBoundaryFuncPointer)();

class DeviceInterface {
public:
    explicit DeviceInterface(int device = 0) : m_Device(device) {}
    ~DeviceInterface() {
        if (m_Ctx) brt_context_destroy(m_Ctx);
    }
    DeviceInterface(const DeviceInterface&) = delete;
    DeviceInterface& operator=(const DeviceInterface&) = delete;

    /// Device initialisation (cudaInit, Core/DeviceInterface.cu:77-102).
    void deviceInit() {
        if (m_Ctx) return;
        int rc = brt_context_create(m_Device, &m_Ctx);
        if (rc != BRT_OK) throw DeviceError("brt_context_create failed (no ROCm device?)");
    }

    brt_context* context() {
        deviceInit();
        return m_Ctx;
    }
    hipStream_t stream() const { return m_Stream; }
    void setStream(hipStream_t s) { m_Stream = s; }

    /// Throws on a non-zero C-ABI status, with the context's message.
    void check(int rc, const char* what) {
        if (rc != BRT_OK)
            throw DeviceError(std::string(what) + ": " + brt_context_last_error(m_Ctx) + " (code " +
                              std::to_string(rc) + ")");
    }

    void allocateArray(void** devPtr, size_t size) {
        deviceInit();
        hipCheck(hipMalloc(devPtr, size), "hipMalloc");
    }
    void freeArray(void* devPtr) {
        if (devPtr) (void)hipFree(devPtr);
    }
    void threadSync() { hipCheck(hipStreamSynchronize(m_Stream), "hipStreamSynchronize"); }
    void copyArrayFromDevice(void* host, const void* device, size_t offset, size_t size) {
        hipCheck(hipMemcpyAsync((char*)host + offset, (const char*)device + offset, size, hipMemcpyDeviceToHost,
                                m_Stream),
                 "hipMemcpy D2H");
        threadSync();
    }
    void copyArrayToDevice(void* device, const void* host, size_t offset, size_t size) {
        hipCheck(hipMemcpyAsync((char*)device + offset, (const char*)host + offset, size, hipMemcpyHostToDevice,
                                m_Stream),
                 "hipMemcpy H2D");
        threadSync();
    }
    void resetArray(void* deviceData, char value, size_t size) {
        hipCheck(hipMemsetAsync(deviceData, value, size, m_Stream), "hipMemset");
    }

    /// kernelCall for custom plug-ins: launches a parameterless kernel on
    /// ceil(total/blockSize) blocks, with optional host hooks before and after
    /// (Core/DeviceInterface.cu:380-458).
    void kernelCall(int totalThreadNum, int blockSize, genericKernelFuncPointer kernelRef,
                    kernelBoundaryFuncPointer beforeKernelRef = nullptr,
                    kernelBoundaryFuncPointer afterKernelRef = nullptr) {
        deviceInit();
        if (totalThreadNum <= 0) return;
        int threads = blockSize < totalThreadNum ? blockSize : totalThreadNum;
        int blocks = (totalThreadNum + threads - 1) / threads;
        if (beforeKernelRef) beforeKernelRef();
        hipCheck(hipLaunchKernel((const void*)kernelRef, dim3(blocks), dim3(threads), nullptr, 0, m_Stream),
                 "hipLaunchKernel");
        if (afterKernelRef) afterKernelRef();
        if (m_HostSimParams.useThreadSync) threadSync();
    }
    void kernelCall(int totalThreadNum, genericKernelFuncPointer kernelRef,
                    kernelBoundaryFuncPointer beforeKernelRef = nullptr,
                    kernelBoundaryFuncPointer afterKernelRef = nullptr) {
        kernelCall(totalThreadNum, m_HostSimParams.commonBlockDim, kernelRef, beforeKernelRef, afterKernelRef);
    }

    SimParams& getHostSimParams() { return m_HostSimParams; }

private:
    int m_Device;
    brt_context* m_Ctx = nullptr;
    hipStream_t m_Stream = nullptr;
    SimParams m_HostSimParams;
};

}  // namespace BehaveRT
