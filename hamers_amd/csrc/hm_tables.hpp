// hm_tables.hpp -- host images of segment work tables (see SegTableT).
#pragma once

#include <cstring>
#include <vector>

#include "hm_device.hpp"

namespace hm {

inline size_t up256(size_t b) { return (b + 255) / 256 * 256; }

// Segment tables packed into one host image: items, then start/cnt/seg.
template <class T>
struct TableImage {
    std::vector<T> items;
    std::vector<int64_t> start, cnt;
    std::vector<int> seg;
    int64_t total = 0;
    void add(int item, int axis, int64_t count) {
        if (count <= 0) return;
        start.push_back(total);
        cnt.push_back(count);
        seg.push_back(item * 4 + axis);
        total += (count + kSegAlign - 1) / kSegAlign * kSegAlign;
    }
    size_t bytes() const {
        return up256(sizeof(T) * items.size()) + up256(sizeof(int64_t) * (start.size() + 1)) +
               up256(sizeof(int64_t) * cnt.size()) + up256(sizeof(int) * seg.size());
    }
    // Writes the image at host pointer h, returns the device view for device base d.
    SegTableT<T> write(char* h, char* d) {
        std::vector<int64_t> st = start;
        st.push_back(total);
        size_t o = 0;
        SegTableT<T> t{};
        std::memcpy(h + o, items.data(), sizeof(T) * items.size());
        t.patches = reinterpret_cast<const T*>(d + o);
        o += up256(sizeof(T) * items.size());
        std::memcpy(h + o, st.data(), sizeof(int64_t) * st.size());
        t.start = reinterpret_cast<const int64_t*>(d + o);
        o += up256(sizeof(int64_t) * st.size());
        if (!cnt.empty()) std::memcpy(h + o, cnt.data(), sizeof(int64_t) * cnt.size());
        t.cnt = reinterpret_cast<const int64_t*>(d + o);
        o += up256(sizeof(int64_t) * cnt.size());
        if (!seg.empty()) std::memcpy(h + o, seg.data(), sizeof(int) * seg.size());
        t.seg = reinterpret_cast<const int*>(d + o);
        t.nseg = int(cnt.size());
        return t;
    }
};

}  // namespace hm
