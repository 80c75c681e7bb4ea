// ZGraphCompat.hpp -- single-process stand-ins for the reference's runtime
// helpers that its application mains call around the estimator path
// (include/NetworkInterface.hpp, include/SharedMemSys.hpp, include/Debug.hpp,
// include/Time.hpp, MPI_Barrier).  One process drives one GPU, so the
// network and NUMA layers have nothing to do: ranks = 1, partition 0.
#ifndef ZGRAPH_AMD_COMPAT_HPP
#define ZGRAPH_AMD_COMPAT_HPP

#include <chrono>
#include <iostream>
#include <map>
#include <string>

namespace zgraph_compat {

class NetworkInterface {
  public:
    static void init_network_interface(int *, char ***) {}
    static NetworkInterface *get_instance()
    {
        static NetworkInterface n;
        return &n;
    }
    int get_partition_id() const { return 0; }
    int get_num_partitions() const { return 1; }
    static void finalize_instance() {}
};

class SharedMemSys {
  public:
    static void init_shared_mem_sys() {}
    static SharedMemSys *get_instance()
    {
        static SharedMemSys s;
        return &s;
    }
    int get_num_sockets() const { return 1; }
    int get_num_threads_per_socket() const { return 1; }
};

class Debug {
  public:
    static Debug *get_instance()
    {
        static Debug d;
        return &d;
    }
    void enter_function(const std::string &) {}
    void leave_function(const std::string &) {}
    template <typename... A>
    void log(const A &...a)
    {
        (std::cerr << ... << a) << std::endl;
    }
    template <typename... A>
    void print(const A &...a)
    {
        (std::cout << ... << a) << std::endl;
    }
};

class Timer {
  public:
    static void timer_start(const std::string &name) { starts()[name] = std::chrono::steady_clock::now(); }
    static double timer_stop(const std::string &name)
    {
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - starts()[name]).count();
        std::cerr << "[timer] " << name << ": " << s << " s" << std::endl;
        totals()[name] += s;
        return s;
    }
    static void report_timers()
    {
        for (const auto &kv : totals()) std::cerr << "[timer] total " << kv.first << ": " << kv.second << " s" << std::endl;
    }

  private:
    static std::map<std::string, std::chrono::steady_clock::time_point> &starts()
    {
        static std::map<std::string, std::chrono::steady_clock::time_point> m;
        return m;
    }
    static std::map<std::string, double> &totals()
    {
        static std::map<std::string, double> m;
        return m;
    }
};

}  // namespace zgraph_compat

using zgraph_compat::Debug;
using zgraph_compat::NetworkInterface;
using zgraph_compat::SharedMemSys;
using zgraph_compat::Timer;

enum { MPI_COMM_WORLD = 0 };
inline int MPI_Barrier(int) { return 0; }

#endif
