// Output-callback fan-out: one consumer thread per registered callback, as
// OCTproEngine's CallbackManager (src/core/callback_manager.h:11-78,
// callback_manager.cpp:6-135).  invokeAll() posts the buffer pointer to
// every consumer queue and returns; each callback runs on its own thread.
// The buffer stays valid until the backend reuses its output slot.
#ifndef OPE_CALLBACK_MANAGER_H
#define OPE_CALLBACK_MANAGER_H

#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "iobuffer.h"

namespace ope {

class CallbackManager {
public:
    using CallbackId = int;
    using OutputCallback = std::function<void(const IOBuffer&)>;

    CallbackManager();
    ~CallbackManager();
    CallbackManager(const CallbackManager&) = delete;
    CallbackManager& operator=(const CallbackManager&) = delete;

    CallbackId addCallback(OutputCallback callback);
    bool removeCallback(CallbackId id);  // blocks until its current call returns
    void clear();
    void invokeAll(const IOBuffer& buffer);
    size_t getCallbackCount() const;

private:
    class Consumer;
    std::vector<std::unique_ptr<Consumer>> consumers_;
    mutable std::mutex mu_;
    std::atomic<CallbackId> next_id_{0};
};

}  // namespace ope

#endif  // OPE_CALLBACK_MANAGER_H
