
#include <thread>
#include <vector>
extern "C" int planted_overflow(int n) {
  std::vector<int> v(n, 1);
  int* p = v.data();
  volatile int i = n;
  return p[i];
}
extern "C" int planted_race(int n) {
  int x = 0;
  std::thread a([&] { for (int i = 0; i < n; ++i) x += 1; });
  std::thread b([&] { for (int i = 0; i < n; ++i) x += 1; });
  a.join();
  b.join();
  return x;
}
