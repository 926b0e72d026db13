// rrs_bowl.cpp -- test harness: the generic Recursive Random Search of the
// planner (fastmulticolsort_amd/csrc/fmcs_rrs.h) on the two bowls of the
// reference's RRS test (tests/rrs_test.cpp:78-100): a 2-D point in the
// square of radius 10 around the origin, neighbours uniform within
// radius * ro of the point, 1000 exploration rounds.
#include <cmath>
#include <cstdint>
#include <vector>

#include "fmcs_rrs.h"

namespace {

class Bowl : public fmcs::rrs::Search<std::vector<double>, double> {
 public:
  Bowl(int which, uint64_t seed) : Search(seed), which_(which) {}

 protected:
  std::vector<double> Sample() override {
    const double x = Dice() * 2 * radius_ - radius_;
    const double y = Dice() * 2 * radius_ - radius_;
    return {x, y};
  }
  double Evaluate(const std::vector<double>& p) override {
    const double x = p[0], y = p[1];
    if (which_ == 1) return (x - 1) * (x - 1) + (y - 2) * (y - 2);
    return x > 0.0 ? (x - 1) * (x - 1) + (y - 2) * (y - 2) : (x + 5) * (x + 5) + (y + 1) * (y + 1) - 1;
  }
  bool Stop() override { return ++rounds_ > 1000; }
  std::vector<double> Neighbor(const std::vector<double>& p, double ro) override {
    const double x = p[0] + (Dice() - 0.5) * radius_ * ro;
    const double y = p[1] + (Dice() - 0.5) * radius_ * ro;
    return {x, y};
  }

 private:
  int which_;
  double radius_ = 10;
  size_t rounds_ = 0;
};

}  // namespace

extern "C" double rrs_bowl(int which, uint64_t seed, double* x, double* y) {
  Bowl b(which, seed);
  b.Optimize();
  *x = b.Best()[0];
  *y = b.Best()[1];
  return b.BestCost();
}
