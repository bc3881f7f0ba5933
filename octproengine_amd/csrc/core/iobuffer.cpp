// IOBuffer and DataType helpers (semantics of OCTproEngine's
// src/core/iobuffer.cpp:46-119 and src/core/types.cpp:5-53): 64-byte aligned,
// zero-initialised host memory.  A backend may instead hand a buffer pinned
// memory through detail::IOBufferAccess, with the matching release function.
#include <cstdlib>
#include <cstring>

#include "iobuffer.h"
#include "iobuffer_access.h"
#include "types.h"

namespace ope {

int getDataTypeBitDepth(DataType type) {
    switch (type) {
        case DataType::UINT8:
        case DataType::INT8: return 8;
        case DataType::UINT16:
        case DataType::INT16: return 16;
        case DataType::UINT32:
        case DataType::INT32:
        case DataType::FLOAT32: return 32;
        case DataType::UINT64:
        case DataType::INT64:
        case DataType::FLOAT64:
        case DataType::COMPLEX_FLOAT32: return 64;
        case DataType::COMPLEX_FLOAT64: return 128;
    }
    return 0;
}

int getDataTypeByteSize(DataType type) { return getDataTypeBitDepth(type) / 8; }

IOBuffer::IOBuffer() : dataPtr(nullptr), sizeInBytes(0), dataType(DataType::UINT8), freeFn(nullptr) {}

IOBuffer::~IOBuffer() { this->releaseMemory(); }

bool IOBuffer::allocateMemory(size_t bytes) {
    this->releaseMemory();
    if (bytes == 0) return true;
    void* p = nullptr;
    if (posix_memalign(&p, 64, bytes) != 0) return false;
    std::memset(p, 0, bytes);
    this->dataPtr = p;
    this->sizeInBytes = bytes;
    this->freeFn = std::free;
    return true;
}

void IOBuffer::releaseMemory() {
    if (this->dataPtr) {
        if (this->freeFn) this->freeFn(this->dataPtr);
        this->dataPtr = nullptr;
        this->sizeInBytes = 0;
        this->freeFn = nullptr;
    }
}

void IOBuffer::setDataType(DataType type) { this->dataType = type; }
IOBuffer::DataType IOBuffer::getDataType() const { return this->dataType; }
int IOBuffer::getBitDepth() const { return getDataTypeBitDepth(this->dataType); }
void* IOBuffer::getDataPointer() { return this->dataPtr; }
const void* IOBuffer::getDataPointer() const { return this->dataPtr; }
size_t IOBuffer::getSizeInBytes() const { return this->sizeInBytes; }

namespace detail {

void IOBufferAccess::adopt(IOBuffer& b, void* ptr, size_t bytes, void (*release)(void*)) {
    b.releaseMemory();
    b.dataPtr = ptr;
    b.sizeInBytes = ptr ? bytes : 0;
    b.freeFn = release;
}

}  // namespace detail
}  // namespace ope
