// fmcs_rrs.h -- Recursive Random Search (Ye & Kalyanaraman, 2003) over any
// point type: the black-box search the plan enumerator runs on massaging
// plans (src/rrs.h:24-185 is the reference's; the steps, the parameters and
// the engine are kept so that a plan enumeration visits the same plans for
// the same seed).  A problem derives from Search and supplies a sampler of
// the whole space, a cost, a neighbour within a shrinking radius and a stop
// rule.  Host C++ only.
#pragma once
#include <cmath>
#include <cstdint>
#include <numeric>
#include <random>
#include <vector>

namespace fmcs {
namespace rrs {

template <class Point, class Cost>
class Search {
 public:
  // p: confidence that the initial samples hit the r-best region; q, v: the
  // same for the exploitation's local samples; c: radius shrink factor; st:
  // the smallest radius.  The engine is the C++ library's minstd_rand0, the
  // default_random_engine of the reference's build (libstdc++).
  explicit Search(uint64_t seed, double p = 0.99, double r = 0.1, double q = 0.99, double v = 0.8,
                  double c = 0.5, double st = 0.001)
      : p_(p), r_(r), q_(q), v_(v), c_(c), st_(st), gen_((std::minstd_rand0::result_type)seed), dist_(0.0, 1.0) {}
  virtual ~Search() {}

  void Optimize() {
    const size_t nsamp = (size_t)std::ceil(std::log(1 - p_) / std::log(1 - r_));
    const size_t l = (size_t)std::ceil(std::log(1 - q_) / std::log(1 - v_));
    // exploration: the best of nsamp samples of the whole space
    best_ = Sample();
    best_cost_ = Evaluate(best_);
    for (size_t i = 1; i < nsamp; ++i) {
      Point x = Sample();
      const Cost cx = Evaluate(x);
      if (cx < best_cost_) {
        best_cost_ = cx;
        best_ = x;
      }
    }
    Point x0 = best_;
    Cost cost0 = best_cost_, threshold = best_cost_;
    std::vector<Cost> history{best_cost_};
    bool exploit = true;
    size_t i = 0;
    while (!Stop()) {
      if (exploit) {
        // exploitation: neighbours of x0 within radius ro; re-align on every
        // improvement, shrink the radius after l failures in a row
        size_t fails = 0;
        Cost fc = Evaluate(x0);
        Point xl = x0;
        double ro = r_;
        while (ro > st_) {
          Point x1 = Neighbor(xl, ro);
          const Cost c1 = Evaluate(x1);
          if (c1 < fc) {
            xl = x1;
            fc = c1;
            fails = 0;
          } else {
            ++fails;
          }
          if (fails == l) {
            ro = c_ * ro;
            fails = 0;
          }
        }
        exploit = false;
        if (fc < best_cost_) {
          best_cost_ = fc;
          best_ = xl;
        }
      }
      // exploration: a new sample; a promising one starts an exploitation
      Point x1 = Sample();
      const Cost c1 = Evaluate(x1);
      if (c1 < cost0) {
        x0 = x1;
        cost0 = c1;
      }
      if (c1 < threshold) exploit = true;
      if (i == nsamp) {  // the threshold follows the mean of the best costs, every nsamp samples
        history.push_back(cost0);
        threshold = std::accumulate(history.begin(), history.end(), Cost(0)) / (Cost)history.size();
        i = 0;
      }
      ++i;
    }
  }

  const Point& Best() const { return best_; }
  Cost BestCost() const { return best_cost_; }

 protected:
  virtual Point Sample() = 0;
  virtual Cost Evaluate(const Point& x) = 0;
  virtual bool Stop() = 0;
  virtual Point Neighbor(const Point& x, double ro) = 0;
  double Dice() { return dist_(gen_); }

 private:
  double p_, r_, q_, v_, c_, st_;
  std::minstd_rand0 gen_;
  std::uniform_real_distribution<double> dist_;
  Point best_{};
  Cost best_cost_{};
};

}  // namespace rrs
}  // namespace fmcs
