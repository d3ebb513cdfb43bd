// DeviceArrayWrapper.h -- a double-buffered per-agent feature (host image + two
// device buffers).
//
// Mirrors BehaveRT::DeviceArrayWrapper<T> (Core/include/DeviceArrayWrapper.h:58-192,
// 198-477): getRead/WriteDeviceArray, swapPosReadWrite, copyArrayToDevice
// (writes the WRITE buffer), copyArrayFromDevice (reads the READ buffer),
// get/setHostArrayElement, reset (clears the WRITE buffer, then swaps).  The GL
// VBO options are dropped.
#pragma once

#include <cstring>
#include <typeinfo>
#include <utility>
#include <vector>

#include "DeviceInterface.h"

namespace BehaveRT {

class InstallableEntity {
public:
    virtual ~InstallableEntity() {}
    void setInstallationKey(int k) { m_InstallationKey = k; }
    int getInstallationKey() const { return m_InstallationKey; }

private:
    int m_InstallationKey = -1;
};

template <class Type>
class DeviceArrayWrapper : public InstallableEntity {
public:
    DeviceArrayWrapper(DeviceInterface* di, int size, int dimension = 1, bool useDoubleBufferingOnDevice = true)
        : m_DeviceInterface(di), m_Size(size), m_Dimension(dimension), m_DoubleBuffer(useDoubleBufferingOnDevice) {
        m_HostArray.resize((size_t)size * dimension);
        m_DeviceArray[0] = m_DeviceArray[1] = nullptr;
        if (getBytesCount() > 0) {
            m_DeviceInterface->allocateArray((void**)&m_DeviceArray[0], getBytesCount());
            if (m_DoubleBuffer) m_DeviceInterface->allocateArray((void**)&m_DeviceArray[1], getBytesCount());
        }
        m_Read = 0;
        m_Write = m_DoubleBuffer ? 1 : 0;
    }
    ~DeviceArrayWrapper() {
        m_DeviceInterface->freeArray(m_DeviceArray[0]);
        if (m_DoubleBuffer) m_DeviceInterface->freeArray(m_DeviceArray[1]);
    }
    DeviceArrayWrapper(const DeviceArrayWrapper&) = delete;
    DeviceArrayWrapper& operator=(const DeviceArrayWrapper&) = delete;

    Type* getReadDeviceArray() { return m_DeviceArray[m_Read]; }
    Type* getWriteDeviceArray() { return m_DeviceArray[m_Write]; }
    size_t getElementBytesCount() const { return sizeof(Type) * (size_t)m_Dimension; }
    size_t getBytesCount() const { return getElementBytesCount() * (size_t)m_Size; }
    Type* getHostArray() { return m_HostArray.data(); }
    int getSize() const { return m_Size; }
    const std::type_info* getType() const { return &typeid(Type); }

    void swapPosReadWrite() {
        if (m_DoubleBuffer) std::swap(m_Read, m_Write);
    }
    void copyArrayToDevice() { copyArrayToDevice(0, m_Size); }
    void copyArrayToDevice(int offsetElements, int numElements) {
        m_DeviceInterface->copyArrayToDevice(m_DeviceArray[m_Write], m_HostArray.data(),
                                             offsetElements * getElementBytesCount(),
                                             numElements * getElementBytesCount());
    }
    void copyArrayFromDevice() { copyArrayFromDevice(0, m_Size); }
    void copyArrayFromDevice(int offsetElements, int numElements) {
        m_DeviceInterface->copyArrayFromDevice(m_HostArray.data(), m_DeviceArray[m_Read],
                                               offsetElements * getElementBytesCount(),
                                               numElements * getElementBytesCount());
    }
    Type getHostArrayElement(int index, bool copyFromDevice = false) {
        if (copyFromDevice) copyArrayFromDevice(index, 1);
        return m_HostArray[(size_t)index * m_Dimension];
    }
    Type* getHostArrayElementRef(int index, bool copyFromDevice = false) {
        if (copyFromDevice) copyArrayFromDevice(index, 1);
        return &m_HostArray[(size_t)index * m_Dimension];
    }
    void setHostArrayElement(int index, const Type* element, bool copyToDevice = false) {
        for (int d = 0; d < m_Dimension; ++d) m_HostArray[(size_t)index * m_Dimension + d] = element[d];
        if (copyToDevice) copyArrayToDevice(index, 1);
    }
    /// Clears the write buffer to `val` bytes and swaps (DeviceArrayWrapper::reset).
    void reset(char val) {
        m_DeviceInterface->resetArray(m_DeviceArray[m_Write], val, getBytesCount());
        swapPosReadWrite();
    }

private:
    DeviceInterface* m_DeviceInterface;
    int m_Size;
    int m_Dimension;
    bool m_DoubleBuffer;
    std::vector<Type> m_HostArray;
    Type* m_DeviceArray[2];
    int m_Read, m_Write;
};

}  // namespace BehaveRT
