/* hash_pair.h -- the bucket element, layout-identical to kstppd/hashinator
 * include/hashinator/hash_pair.h:29-75 ({T first; U second;}). */
#pragma once
#include "../common.h"

namespace Hashinator {

template <typename T, typename U>
struct hash_pair {
   T first;
   U second;
   HASHINATOR_HOSTDEVICE hash_pair() : first(T()), second(U()) {}
   HASHINATOR_HOSTDEVICE hash_pair(const T& f, const U& s) : first(f), second(s) {}
   HASHINATOR_HOSTDEVICE inline bool operator==(const hash_pair& y) const { return first == y.first && second == y.second; }
   HASHINATOR_HOSTDEVICE inline bool operator!=(const hash_pair& y) const { return !(*this == y); }
};

template <class T, class U>
HASHINATOR_HOSTDEVICE inline hash_pair<T, U> make_pair(T x, U y) {
   return hash_pair<T, U>(x, y);
}

} // namespace Hashinator
