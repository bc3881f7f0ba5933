#include "callback_manager.h"

#include <algorithm>

namespace ope {

class CallbackManager::Consumer {
public:
    Consumer(CallbackId id, OutputCallback cb) : id(id), cb_(std::move(cb)) {
        thread_ = std::thread([this] { this->run(); });
    }
    ~Consumer() { stop(); }

    void post(const IOBuffer* b) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            queue_.push_back(b);
        }
        cv_.notify_one();
    }

    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (stopping_) return;
            stopping_ = true;
        }
        cv_.notify_one();
        if (thread_.joinable()) thread_.join();
    }

    const CallbackId id;

private:
    void run() {
        for (;;) {
            const IOBuffer* b = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stopping_ || !queue_.empty(); });
                if (stopping_) return;
                b = queue_.front();
                queue_.pop_front();
            }
            try {
                cb_(*b);
            } catch (...) {
                // a failing consumer must not take the pipeline down
                // (callback_manager.cpp:41-48 swallows the same way)
            }
        }
    }

    OutputCallback cb_;
    std::thread thread_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<const IOBuffer*> queue_;
    bool stopping_ = false;
};

CallbackManager::CallbackManager() = default;

CallbackManager::~CallbackManager() { clear(); }

CallbackManager::CallbackId CallbackManager::addCallback(OutputCallback callback) {
    std::lock_guard<std::mutex> lk(mu_);
    const CallbackId id = next_id_++;
    consumers_.push_back(std::make_unique<Consumer>(id, std::move(callback)));
    return id;
}

bool CallbackManager::removeCallback(CallbackId id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = std::find_if(consumers_.begin(), consumers_.end(),
                           [id](const std::unique_ptr<Consumer>& c) { return c->id == id; });
    if (it == consumers_.end()) return false;
    (*it)->stop();
    consumers_.erase(it);
    return true;
}

void CallbackManager::clear() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& c : consumers_) c->stop();
    consumers_.clear();
}

void CallbackManager::invokeAll(const IOBuffer& buffer) {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& c : consumers_) c->post(&buffer);
}

size_t CallbackManager::getCallbackCount() const {
    std::lock_guard<std::mutex> lk(mu_);
    return consumers_.size();
}

}  // namespace ope
