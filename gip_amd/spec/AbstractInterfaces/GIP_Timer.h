// Timer abstraction (interface of GIP_Timer.h:6-26 in the reference):
// a channel is started and stopped; local and global (all-partition)
// elapsed times in seconds are accumulated.
#ifndef _gip_timer_
#define _gip_timer_

class GIP_Timer {
   public:
    GIP_Timer() {}
    virtual ~GIP_Timer() {}
    virtual void startChannel() {}
    virtual void stopChannel() {}
    virtual void stopAllChannels() {}
    float getTimeL() { return tm_l; }
    float getSumL() { return sum_l; }
    float getTimeG() { return tm_g; }
    float getSumG() { return sum_g; }
    float getTimeAvg() { return cont ? sum_g / cont : 0.0f; }

   protected:
    float tm_l = 0, sum_l = 0, tm_g = 0, sum_g = 0;
    int cont = 0;
};

#endif
