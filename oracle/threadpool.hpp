// TEST INFRASTRUCTURE (oracle).  Work-queue thread pool with the interface of VulkanExp's
// ThreadPool (src/ThreadPool.hpp:10-75, src/ThreadPool.cpp:5-28): a process-wide instance,
// `queue(fn) -> std::future<void>`, and a `TaskQueue` whose `start` enqueues and `wait` joins.
// Unlike the reference, the instance can be resized (the CPU baseline states its thread count),
// and `parallel_for` splits an index range into fixed-size chunks.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <future>
#include <mutex>
#include <thread>
#include <vector>

namespace orc {

class ThreadPool {
  public:
	class TaskQueue {
	  public:
		~TaskQueue() { wait(); }
		const std::future<void>& start(std::function<void()>&& fn) {
			pending_.push_back(ThreadPool::GetInstance().queue(std::move(fn)));
			return pending_.back();
		}
		void wait() {
			for(auto& f : pending_) f.get();
			pending_.clear();
		}

	  private:
		std::vector<std::future<void>> pending_;
	};

	explicit ThreadPool(uint32_t threads = default_threads()) { start(threads); }
	ThreadPool(const ThreadPool&) = delete;
	ThreadPool& operator=(const ThreadPool&) = delete;
	~ThreadPool() { stop(); }

	static uint32_t default_threads() {
		uint32_t n = std::thread::hardware_concurrency();
		return n == 0 ? 1 : n;
	}

	static ThreadPool& GetInstance() {
		static ThreadPool pool;
		return pool;
	}

	uint32_t size() const { return static_cast<uint32_t>(workers_.size()); }

	// Stop all workers and start `threads` new ones (0 = hardware_concurrency()).
	void resize(uint32_t threads) {
		stop();
		start(threads == 0 ? default_threads() : threads);
	}

	std::future<void> queue(std::function<void()>&& fn) {
		std::packaged_task<void()> task(std::move(fn));
		auto fut = task.get_future();
		{
			std::lock_guard<std::mutex> lk(mu_);
			jobs_.push_back(std::move(task));
		}
		cv_.notify_one();
		return fut;
	}

	// Run body(begin, end) over [0, n) in chunks of `chunk`, on the pool, and wait.
	template <class F>
	static void parallel_for(uint64_t n, uint64_t chunk, F&& body) {
		if(n == 0) return;
		chunk = std::max<uint64_t>(1, chunk);
		ThreadPool& pool = GetInstance();
		if(pool.size() <= 1 || n <= chunk) {
			body(uint64_t(0), n);
			return;
		}
		TaskQueue q;
		for(uint64_t b = 0; b < n; b += chunk) {
			uint64_t e = std::min(n, b + chunk);
			q.start([&body, b, e] { body(b, e); });
		}
		q.wait();
	}

  private:
	std::vector<std::thread> workers_;
	std::deque<std::packaged_task<void()>> jobs_;
	std::mutex mu_;
	std::condition_variable cv_;

	void start(uint32_t threads) {
		workers_.reserve(threads);
		for(uint32_t i = 0; i < threads; ++i) workers_.emplace_back([this] { loop(); });
	}

	void stop() {
		{
			std::lock_guard<std::mutex> lk(mu_);
			for(size_t i = 0; i < workers_.size(); ++i) jobs_.emplace_back();  // empty task = exit
		}
		cv_.notify_all();
		for(auto& t : workers_) t.join();
		workers_.clear();
	}

	void loop() {
		for(;;) {
			std::packaged_task<void()> job;
			{
				std::unique_lock<std::mutex> lk(mu_);
				cv_.wait(lk, [this] { return !jobs_.empty(); });
				job = std::move(jobs_.front());
				jobs_.pop_front();
			}
			if(!job.valid()) return;
			job();
		}
	}
};

}  // namespace orc
