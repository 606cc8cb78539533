#include "GIP_TimerHIP.h"

GIP_TimerHIP::GIP_TimerHIP(GIP_Arch *node) : ctx_(GIP_ArchHIP::of(node)->ctx()) {
    gip_check(gip_event_create(ctx_, &start_), "event create");
    gip_check(gip_event_create(ctx_, &stop_), "event create");
}

GIP_TimerHIP::~GIP_TimerHIP() {
    gip_event_destroy(start_);
    gip_event_destroy(stop_);
}

void GIP_TimerHIP::startChannel() { gip_check(gip_event_record(ctx_, start_), "event record"); }

void GIP_TimerHIP::stopChannel() {
    gip_check(gip_event_record(ctx_, stop_), "event record");
    float ms = 0.0f;
    gip_check(gip_event_elapsed_ms(start_, stop_, &ms), "event elapsed");
    tm_l = ms * 1e-3f;
    sum_l += tm_l;
    cont += 1;
}

void GIP_TimerHIP::stopAllChannels() {
    stopChannel();
    tm_g = tm_l;
    sum_g += tm_g;
}
