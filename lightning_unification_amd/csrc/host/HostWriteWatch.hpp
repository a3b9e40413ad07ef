// Detection of host writes into a caller-owned buffer that has a device copy.
//
// Lightning-Qubit's StateVectorC* is a view of the caller's numpy buffer
// (StateVectorLQubitRaw.hpp:78-87): whatever Python wrote into the buffer is
// what the next call sees.  The drop-in keeps the amplitudes in HBM instead,
// so before it reuses a device copy it must know whether the host buffer was
// written since the two last agreed.  A WriteWatch answers that exactly and
// at no cost while nothing is written:
//   - the pages lying wholly inside the buffer are made read-only; the first
//     write to any of them faults, the SIGSEGV handler below records the
//     fault against the watch, makes the range writable again and returns,
//     and the write then completes as if nothing had happened;
//   - the bytes of the buffer that share a page with other memory (at most
//     one partial page at each end) are kept as a snapshot and compared.
// Faults outside every watched range go to the handler that was installed
// before this one (Python's faulthandler, or the default action).
//
// Limits: writes made by the kernel on the process's behalf (read(2) into the
// buffer) fail with EFAULT instead of faulting, and a write by another
// process through shared memory is not seen.  Neither happens to a device's
// state buffer.
#pragma once

#include <signal.h>
#include <sys/mman.h>
#include <unistd.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace Pennylane::LightningAMD::hostwatch {

constexpr int kSlots = 256;

struct Slot {
    // page-aligned interior range [begin, end); begin == 0: slot unused
    std::atomic<uintptr_t> begin{0}, end{0};
    std::atomic<int> faulted{0};
};

inline Slot *slots() {
    static Slot s[kSlots];
    return s;
}

inline struct sigaction &previous_action() {
    static struct sigaction old {};
    return old;
}

// Protection is per page, not per watch: when the pages [b, e) become
// writable, every watch over any of them is marked written.
inline void mark_overlapping(uintptr_t b, uintptr_t e) {
    Slot *s = slots();
    for (int i = 0; i < kSlots; i++) {
        const uintptr_t sb = s[i].begin.load(std::memory_order_acquire);
        const uintptr_t se = s[i].end.load(std::memory_order_acquire);
        if (sb != 0 && sb < e && b < se) {
            s[i].faulted.store(1, std::memory_order_release);
        }
    }
}

inline void on_segv(int sig, siginfo_t *si, void *ctx) {
    const auto a = reinterpret_cast<uintptr_t>(si->si_addr);
    if (si->si_code == SEGV_ACCERR) {
        Slot *s = slots();
        bool hit = false;
        for (int i = 0; i < kSlots; i++) {
            const uintptr_t b = s[i].begin.load(std::memory_order_acquire);
            const uintptr_t e = s[i].end.load(std::memory_order_acquire);
            if (b != 0 && a >= b && a < e) {
                // mprotect is a plain system call; it is safe in a handler
                (void)mprotect(reinterpret_cast<void *>(b), e - b, PROT_READ | PROT_WRITE);
                mark_overlapping(b, e);
                hit = true;
            }
        }
        if (hit) {
            return;
        }
    }
    const struct sigaction &old = previous_action();
    if (old.sa_flags & SA_SIGINFO) {
        if (old.sa_sigaction) {
            old.sa_sigaction(sig, si, ctx);
        }
        return;
    }
    if (old.sa_handler == SIG_IGN) {
        return;
    }
    if (old.sa_handler == SIG_DFL || old.sa_handler == nullptr) {
        // re-executing the faulting instruction takes the default action
        signal(sig, SIG_DFL);
        return;
    }
    old.sa_handler(sig);
}

inline std::mutex &install_mutex() {
    static std::mutex m;
    return m;
}

inline void install_handler() {
    static bool done = false;
    std::lock_guard<std::mutex> lk(install_mutex());
    if (done) {
        return;
    }
    struct sigaction sa {};
    sa.sa_sigaction = on_segv;
    sigemptyset(&sa.sa_mask);
    sa.sa_flags = SA_SIGINFO | SA_RESTART | SA_ONSTACK;
    if (sigaction(SIGSEGV, &sa, &previous_action()) != 0) {
        throw std::runtime_error("cannot install the host write-watch handler");
    }
    done = true;
}

inline size_t page_size() {
    static const size_t p = static_cast<size_t>(sysconf(_SC_PAGESIZE));
    return p;
}

/* One watched buffer.  watch() starts watching (the buffer's current bytes
 * become the reference), written() tells whether anything was written since,
 * release() stops watching and leaves the buffer writable. */
class WriteWatch {
  public:
    WriteWatch() = default;
    WriteWatch(const WriteWatch &) = delete;
    WriteWatch &operator=(const WriteWatch &) = delete;
    ~WriteWatch() { release(); }

    void watch(const void *data, size_t nbytes) {
        release();
        base_ = reinterpret_cast<uintptr_t>(data);
        len_ = nbytes;
        const size_t pg = page_size();
        uintptr_t ib = (base_ + pg - 1) & ~(uintptr_t(pg) - 1);
        uintptr_t ie = (base_ + len_) & ~(uintptr_t(pg) - 1);
        if (ie <= ib) {
            ib = ie = 0; // no whole page inside: everything is an edge
        }
        head_.assign(reinterpret_cast<const char *>(base_),
                     reinterpret_cast<const char *>(ib ? ib : base_ + len_));
        tail_.assign(reinterpret_cast<const char *>(ie ? ie : base_ + len_),
                     reinterpret_cast<const char *>(base_ + len_));
        if (ib) {
            install_handler();
            slot_ = claim(ib, ie);
            if (slot_ < 0) {
                // every slot busy: fall back to treating the buffer as written
                // on every check (a re-upload, never a stale read)
                overflow_ = true;
            } else if (mprotect(reinterpret_cast<void *>(ib), ie - ib, PROT_READ) != 0) {
                drop_slot();
                overflow_ = true;
            }
        }
        ib_ = ib;
        ie_ = ie;
        active_ = true;
    }

    [[nodiscard]] bool active() const { return active_; }

    [[nodiscard]] bool written() const {
        if (!active_) {
            return true;
        }
        if (overflow_) {
            return true;
        }
        if (slot_ >= 0 && slots()[slot_].faulted.load(std::memory_order_acquire)) {
            return true;
        }
        const char *p = reinterpret_cast<const char *>(base_);
        if (!head_.empty() && std::memcmp(p, head_.data(), head_.size()) != 0) {
            return true;
        }
        const char *t = reinterpret_cast<const char *>(ie_ ? ie_ : base_ + len_);
        return !tail_.empty() && std::memcmp(t, tail_.data(), tail_.size()) != 0;
    }

    void release() {
        if (!active_) {
            return;
        }
        if (slot_ >= 0) {
            (void)mprotect(reinterpret_cast<void *>(ib_), ie_ - ib_, PROT_READ | PROT_WRITE);
            mark_overlapping(ib_, ie_);
            drop_slot();
        }
        active_ = false;
        overflow_ = false;
        head_.clear();
        tail_.clear();
    }

  private:
    static int claim(uintptr_t b, uintptr_t e) {
        std::lock_guard<std::mutex> lk(install_mutex());
        Slot *s = slots();
        for (int i = 0; i < kSlots; i++) {
            if (s[i].begin.load() == 0) {
                s[i].faulted.store(0);
                s[i].end.store(e, std::memory_order_release);
                s[i].begin.store(b, std::memory_order_release);
                return i;
            }
        }
        return -1;
    }
    void drop_slot() {
        std::lock_guard<std::mutex> lk(install_mutex());
        slots()[slot_].begin.store(0, std::memory_order_release);
        slot_ = -1;
    }

    uintptr_t base_ = 0, ib_ = 0, ie_ = 0;
    size_t len_ = 0;
    int slot_ = -1;
    bool active_ = false, overflow_ = false;
    std::vector<char> head_, tail_;
};

} // namespace Pennylane::LightningAMD::hostwatch
