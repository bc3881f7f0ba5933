// IOBuffer and DataType helpers (semantics of OCTproEngine's
// src/core/iobuffer.cpp:46-119 and src/core/types.cpp:5-53): 64-byte aligned,
// zero-initialised host memory.  A backend may instead hand a buffer pinned
// memory through detail::IOBufferAccess, with the matching release function.
// IOBuffer keeps the reference's three members (include/iobuffer.h:30-33),
// so the release functions of adopted memory live in a registry keyed by
// the data pointer rather than in the object.
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>

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

namespace {

// Release functions of memory adopted through detail::IOBufferAccess.
std::mutex g_adopted_mu;
std::unordered_map<void*, void (*)(void*)>& adopted() {
    static std::unordered_map<void*, void (*)(void*)> m;
    return m;
}

}  // namespace

IOBuffer::IOBuffer() : dataPtr(nullptr), sizeInBytes(0), dataType(DataType::UINT8) {}

IOBuffer::~IOBuffer() { this->releaseMemory(); }

bool IOBuffer::allocateMemory(size_t bytes) {
    this->releaseMemory();
    if (bytes == 0) return true;
    void* p = nullptr;
    if (posix_memalign(&p, 64, bytes) != 0) return false;
    std::memset(p, 0, bytes);
    this->dataPtr = p;
    this->sizeInBytes = bytes;
    return true;
}

void IOBuffer::releaseMemory() {
    if (this->dataPtr) {
        void (*release)(void*) = std::free;
        {
            std::lock_guard<std::mutex> lk(g_adopted_mu);
            auto it = adopted().find(this->dataPtr);
            if (it != adopted().end()) {
                release = it->second;
                adopted().erase(it);
            }
        }
        if (release) release(this->dataPtr);
        this->dataPtr = nullptr;
        this->sizeInBytes = 0;
    }
}

void IOBuffer::setDataType(DataType type) { this->dataType = type; }
IOBuffer::DataType IOBuffer::getDataType() const { return this->dataType; }
int IOBuffer::getBitDepth() const { return getDataTypeBitDepth(this->dataType); }
void* IOBuffer::getDataPointer() { return this->dataPtr; }
const void* IOBuffer::getDataPointer() const { return this->dataPtr; }
size_t IOBuffer::getSizeInBytes() const { return this->sizeInBytes; }

namespace detail {

// The reference's IOBuffer members (include/iobuffer.h:30-33), in order.
struct ReferenceIOBufferLayout {
    void* dataPtr;
    size_t sizeInBytes;
    DataType dataType;
};

void IOBufferAccess::adopt(IOBuffer& b, void* ptr, size_t bytes, void (*release)(void*)) {
    static_assert(sizeof(IOBuffer) == sizeof(ReferenceIOBufferLayout), "IOBuffer size differs from the reference");
    static_assert(alignof(IOBuffer) == alignof(ReferenceIOBufferLayout), "IOBuffer alignment differs");
    static_assert(offsetof(IOBuffer, dataPtr) == offsetof(ReferenceIOBufferLayout, dataPtr), "dataPtr offset");
    static_assert(offsetof(IOBuffer, sizeInBytes) == offsetof(ReferenceIOBufferLayout, sizeInBytes),
                  "sizeInBytes offset");
    static_assert(offsetof(IOBuffer, dataType) == offsetof(ReferenceIOBufferLayout, dataType), "dataType offset");
    b.releaseMemory();
    if (ptr) {
        std::lock_guard<std::mutex> lk(g_adopted_mu);
        adopted()[ptr] = release;
    }
    b.dataPtr = ptr;
    b.sizeInBytes = ptr ? bytes : 0;
}

}  // namespace detail
}  // namespace ope
