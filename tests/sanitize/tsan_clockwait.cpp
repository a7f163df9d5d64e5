// ThreadSanitizer build only (test infrastructure).  libstdc++ 11 implements
// std::condition_variable::wait_for / wait_until on the steady clock with pthread_cond_clockwait,
// which GCC 11's ThreadSanitizer runtime does not intercept: it then never sees the wait release
// and retake the mutex and reports a "double lock" on the next lock of it.  This definition, which
// the executable's calls bind to, waits through the intercepted pthread_cond_timedwait on the
// equivalent CLOCK_REALTIME deadline, so TSan follows the mutex.
#include <pthread.h>
#include <time.h>

extern "C" int pthread_cond_clockwait(pthread_cond_t* cond, pthread_mutex_t* mutex, clockid_t clock,
                                      const struct timespec* abstime) {
  timespec now_clk, now_rt;
  clock_gettime(clock, &now_clk);
  clock_gettime(CLOCK_REALTIME, &now_rt);
  long long ns = (long long)(abstime->tv_sec - now_clk.tv_sec) * 1000000000LL + (abstime->tv_nsec - now_clk.tv_nsec);
  if (ns < 0) ns = 0;
  long long t = (long long)now_rt.tv_nsec + ns;
  timespec rt;
  rt.tv_sec = now_rt.tv_sec + (time_t)(t / 1000000000LL);
  rt.tv_nsec = (long)(t % 1000000000LL);
  return pthread_cond_timedwait(cond, mutex, &rt);
}
