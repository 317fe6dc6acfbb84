// re_algorithm.hpp -- algorithm state and dirty-tracked parameters of the resaxslib
// surface (resaxslib/include/re_algorithm.hpp:26-95).
#ifndef RESAXS_AMD_RE_ALGORITHM_HPP
#define RESAXS_AMD_RE_ALGORITHM_HPP

#include "resaxs.hpp"

namespace resaxs {
namespace algorithm {

// Base of every algorithm object.  Not reference counted: the factories return
// objects allocated with new that the caller deletes.
struct algorithm {
protected:
    enum algorithm_state { alg_blank = 0, alg_initialized, alg_program_loaded, alg_args_set, alg_computing };
    algorithm_state state = alg_blank;

public:
    virtual ~algorithm() = default;

    bool initialized() const { return state != alg_blank; }
    bool program_loaded() const { return state >= alg_program_loaded; }
    bool args_set() const { return state >= alg_args_set; }
    bool computing() const { return state == alg_computing; }

    void initialize() { state = alg_initialized; }
};

// Parameters that remember whether they changed since the last calculation.
class params_base {
public:
    virtual ~params_base() = default;

    bool is_dirty() const { return dirty_; }
    void clear_dirty() { dirty_ = false; }

private:
    bool dirty_ = false;

protected:
    // called once on each clean -> dirty transition
    virtual void on_dirty_set() {}

    void set_dirty()
    {
        if (!dirty_) {
            dirty_ = true;
            on_dirty_set();
        }
    }
};

}  // namespace algorithm
}  // namespace resaxs

#endif
