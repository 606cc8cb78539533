// HIP timer specification: a channel is a pair of events recorded on the
// node's stream (GIP_TimerCuda.cpp:3-32 uses CUDA events the same way);
// elapsed times are in seconds.  With one partition the global time is the
// local one.
#ifndef _gip_timerhip_
#define _gip_timerhip_

#include "GIP_ArchHIP.h"
#include "../../AbstractInterfaces/GIP_Timer.h"

class GIP_TimerHIP : public GIP_Timer {
   public:
    explicit GIP_TimerHIP(GIP_Arch *node);
    ~GIP_TimerHIP();
    GIP_TimerHIP(const GIP_TimerHIP &) = delete;
    GIP_TimerHIP &operator=(const GIP_TimerHIP &) = delete;
    void startChannel();
    void stopChannel();
    void stopAllChannels();

   protected:
    gip_ctx *ctx_;
    gip_event *start_ = nullptr, *stop_ = nullptr;
};

#endif
