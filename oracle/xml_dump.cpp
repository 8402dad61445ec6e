// xml_dump.cpp - TEST INFRASTRUCTURE ONLY.  The product library never links,
// loads or calls it.
//
// Prints the scenes the reference's own XML reader finds in a scene file:
// TinyXML 2.6.2 as the reference vendors it (source/external/tinyxml*.cpp,
// built from those sources by oracle/Makefile into oracle/_ref/), walked
// the way cSPH::LoadScenes walks it (source/SPH/SPH_Scenes.cpp:56-71): the
// root element's <Scene> children in order.  For each scene, one JSON line
// with its attributes in order and the attributes of its <Emitter> and
// <Accel> children (the children Scene::_FromXML reads,
// source/SPH/Scene_Load.cpp:48-62).  The first line reports whether the
// reader met an error and how many scenes it kept.
#include <cstdio>
#include <string>

#include "tinyxml.h"

static std::string js(const char* s) {
    std::string o = "\"";
    for (; *s; ++s) {
        const unsigned char c = (unsigned char)*s;
        if (c == '"' || c == '\\') {
            o += '\\';
            o += (char)c;
        } else if (c < 0x20) {
            char b[8];
            std::snprintf(b, sizeof(b), "\\u%04x", c);
            o += b;
        } else {
            o += (char)c;
        }
    }
    return o + "\"";
}

static std::string attrs(const TiXmlElement* e) {
    std::string o = "[";
    for (const TiXmlAttribute* a = e->FirstAttribute(); a; a = a->Next()) {
        if (o.size() > 1) o += ",";
        o += "[" + js(a->Name()) + "," + js(a->Value()) + "]";
    }
    return o + "]";
}

static std::string kids(const TiXmlElement* e, const char* name) {
    std::string o = "[";
    for (const TiXmlElement* k = e->FirstChildElement(name); k; k = k->NextSiblingElement(name)) {
        if (o.size() > 1) o += ",";
        o += attrs(k);
    }
    return o + "]";
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s SCENES.xml\n", argv[0]);
        return 2;
    }
    TiXmlDocument file;
    const bool ok = file.LoadFile(argv[1]);
    const TiXmlElement* root = file.RootElement();
    int count = 0;
    for (const TiXmlElement* s = root ? root->FirstChildElement("Scene") : nullptr; s;
         s = s->NextSiblingElement("Scene"))
        ++count;
    std::printf("{\"loaded\": %s, \"error\": %s, \"root\": %s, \"count\": %d}\n", ok ? "true" : "false",
                js(file.Error() ? file.ErrorDesc() : "").c_str(), js(root ? root->Value() : "").c_str(), count);
    for (const TiXmlElement* s = root ? root->FirstChildElement("Scene") : nullptr; s;
         s = s->NextSiblingElement("Scene"))
        std::printf("{\"attrs\": %s, \"Emitter\": %s, \"Accel\": %s}\n", attrs(s).c_str(),
                    kids(s, "Emitter").c_str(), kids(s, "Accel").c_str());
    return 0;
}
